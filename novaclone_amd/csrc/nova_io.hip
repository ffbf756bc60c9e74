// nova_io.hip -- output and restart: EnSight Gold files, the run clock and
// FP64 checkpoints.
//
// Replaces Field::write/read (src/Field/Field.hxx:403-491), Mesh::write,
// writeCase and writeRestartInfo (src/Mesh/Mesh.cpp:224-369) and the Time
// clock (src/Time/Time.cpp:23-139).  The reference converts node by node on
// the host and writes with one stream call per float.  Here:
//
//   * the change of layout (stored k-fastest FP64 with no periodic image ->
//     EnSight component-major, i fastest, float32, N+1 nodes a direction)
//     is one LDS-tiled transpose kernel on the solver's stream into a device
//     staging slot;
//   * the device-to-host copy of the slot runs on the io object's own
//     stream, so the solver's next steps overlap it;
//   * a worker thread waits for the copy and writes the file; two staging
//     slots let a second field be packed while the first is written.
//
// Files are written in the order they are queued.  nova_io_wait drains the
// queue and reports the first error the worker met.
#include "nova_internal.hpp"

#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <mutex>
#include <sstream>
#include <thread>

struct nova_time {
    std::vector<double> steps;
    double dt = 0.0, cur = 0.0, end = 0.0;
    int write_every = 0;
    int time_step = 0;
    int file_num = 0;
    bool timebased = false;
    double interval = 0.0;
    bool write = false, log = true;
};

namespace nova {
double mesh_z_of(const nova_ctx *ctx, int kp);
}

namespace {

struct IoSlot {
    void *dev = nullptr;
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool busy = false;
};

// One file: a header and either a staging slot's bytes (after its copy
// event) or host bytes.
struct IoJob {
    std::string file;
    std::string head;
    int slot = -1;
    size_t bytes = 0;
    bool text = false;
};

}  // namespace

struct nova_io {
    nova_ctx *ctx = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev_in = nullptr;
    IoSlot slot[2];
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<IoJob> q;
    int pending = 0;
    bool stop = false;
    int err_code = 0;
    std::string err;
};

namespace nova {
namespace {

constexpr int kTile = 64;     // nodes per tile side
constexpr int kTileRows = 4;  // thread rows of a 256-thread block

// EnSight node counts per direction: N + 1 (the periodic image included).
struct EnsDims {
    int e[3];
};

// F (stored, FP64, k fastest) -> out (component-major, i fastest, float32,
// N+1 nodes a direction; the periodic image node N is node 0).  Block:
// 64 x 4 threads, one 64 (i) x 64 (k) tile of one (component, j) plane.
__global__ void __launch_bounds__(kTile *kTileRows) k_ens_pack(Grid g, EnsDims d, const double *__restrict__ F,
                                                                float *__restrict__ out)
{
    __shared__ float tile[kTile][kTile + 1];
    const int ntk = (d.e[2] + kTile - 1) / kTile;
    const int k0 = (blockIdx.x % ntk) * kTile, i0 = (blockIdx.x / ntk) * kTile;
    const int j = blockIdx.y, c = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int js = j < g.s[1] ? j : j - g.s[1];
    const double *Fc = F + (int64_t)c * g.np;
    {
        const int k = k0 + tx;
        const int ks = k < g.s[2] ? k : k - g.s[2];
        for (int r = ty; r < kTile; r += kTileRows) {
            const int i = i0 + r;
            if (i < d.e[0] && k < d.e[2]) {
                const int is = i < g.s[0] ? i : i - g.s[0];
                tile[r][tx] = (float)Fc[((int64_t)is * g.s[1] + js) * g.s[2] + ks];
            }
        }
    }
    __syncthreads();
    const int i = i0 + tx;
    for (int r = ty; r < kTile; r += kTileRows) {
        const int k = k0 + r;
        if (i < d.e[0] && k < d.e[2]) out[(((int64_t)c * d.e[2] + k) * d.e[1] + j) * d.e[0] + i] = tile[tx][r];
    }
}

// The inverse for the stored nodes (the image entries of the file are not
// read back): in (EnSight order, float32) -> F (stored, FP64).
__global__ void __launch_bounds__(kTile *kTileRows) k_ens_unpack(Grid g, EnsDims d, const float *__restrict__ in,
                                                                  double *__restrict__ F)
{
    __shared__ float tile[kTile][kTile + 1];
    const int ntk = (g.s[2] + kTile - 1) / kTile;
    const int k0 = (blockIdx.x % ntk) * kTile, i0 = (blockIdx.x / ntk) * kTile;
    const int j = blockIdx.y, c = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    {
        const int i = i0 + tx;
        for (int r = ty; r < kTile; r += kTileRows) {
            const int k = k0 + r;
            if (i < g.s[0] && k < g.s[2]) tile[tx][r] = in[(((int64_t)c * d.e[2] + k) * d.e[1] + j) * d.e[0] + i];
        }
    }
    __syncthreads();
    const int k = k0 + tx;
    double *Fc = F + (int64_t)c * g.np;
    for (int r = ty; r < kTile; r += kTileRows) {
        const int i = i0 + r;
        if (i < g.s[0] && k < g.s[2]) Fc[((int64_t)i * g.s[1] + j) * g.s[2] + k] = (double)tile[r][tx];
    }
}

EnsDims ens_dims(const nova_ctx *ctx)
{
    EnsDims d;
    for (int a = 0; a < 3; a++) d.e[a] = ctx->g.n[a] + 1;
    return d;
}

int64_t ens_nodes(const EnsDims &d) { return (int64_t)d.e[0] * d.e[1] * d.e[2]; }

// An 80-byte EnSight string: the text, then NULs.  (The reference writes 80
// bytes from a shorter temporary string; only the text and its first NUL
// are defined there.)
void put80(std::string &s, const char *txt)
{
    char b[80];
    std::memset(b, 0, sizeof(b));
    std::strncpy(b, txt, 79);
    s.append(b, 80);
}

template <class T>
void put(std::string &s, T v)
{
    s.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

std::string num4(int v)
{
    char b[16];
    std::snprintf(b, sizeof(b), "%04d", v);
    return b;
}

std::string num5(int v)
{
    char b[16];
    std::snprintf(b, sizeof(b), "%05d", v);
    return b;
}

// One rank: the reference's processor number is 0.
constexpr int kProcNo = 0;

void worker_main(nova_io *io)
{
    hipSetDevice(io->ctx->device);
    for (;;) {
        IoJob job;
        {
            std::unique_lock<std::mutex> lk(io->mu);
            io->cv.wait(lk, [&] { return io->stop || !io->q.empty(); });
            if (io->q.empty()) return;
            job = std::move(io->q.front());
            io->q.pop_front();
        }
        int code = 0;
        std::string msg;
        const void *data = nullptr;
        if (job.slot >= 0) {
            const hipError_t e = hipEventSynchronize(io->slot[job.slot].ev);
            if (e != hipSuccess) {
                code = NOVA_ERR_HIP;
                msg = std::string("device-to-host copy: ") + hipGetErrorString(e);
            }
            data = io->slot[job.slot].host;
        }
        if (!code) {
            FILE *f = std::fopen(job.file.c_str(), job.text ? "w" : "wb");
            if (!f) {
                code = NOVA_ERR_IO;
                msg = "cannot open " + job.file;
            } else {
                bool ok = std::fwrite(job.head.data(), 1, job.head.size(), f) == job.head.size();
                if (ok && job.bytes) ok = std::fwrite(data, 1, job.bytes, f) == job.bytes;
                if (std::fclose(f) != 0) ok = false;
                if (!ok) {
                    code = NOVA_ERR_IO;
                    msg = "write failed: " + job.file;
                }
            }
        }
        std::lock_guard<std::mutex> lk(io->mu);
        if (job.slot >= 0) io->slot[job.slot].busy = false;
        if (code && !io->err_code) {
            io->err_code = code;
            io->err = msg;
        }
        io->pending--;
        io->cv.notify_all();
    }
}

void enqueue(nova_io *io, IoJob &&job)
{
    std::lock_guard<std::mutex> lk(io->mu);
    io->q.push_back(std::move(job));
    io->pending++;
    io->cv.notify_all();
}

// A free staging slot of at least `bytes` (waits for the worker).
int acquire_slot(nova_io *io, size_t bytes, int &out)
{
    std::unique_lock<std::mutex> lk(io->mu);
    io->cv.wait(lk, [&] { return !io->slot[0].busy || !io->slot[1].busy; });
    const int s = !io->slot[0].busy ? 0 : 1;
    IoSlot &sl = io->slot[s];
    if (sl.cap < bytes) {
        if (sl.dev) NOVA_HIP(hipFree(sl.dev));
        if (sl.host) NOVA_HIP(hipHostFree(sl.host));
        sl.dev = sl.host = nullptr;
        sl.cap = 0;
        NOVA_HIP(hipMalloc(&sl.dev, bytes));
        NOVA_HIP(hipHostMalloc(&sl.host, bytes, hipHostMallocDefault));
        sl.cap = bytes;
    }
    sl.busy = true;
    out = s;
    return NOVA_OK;
}

void release_slot(nova_io *io, int s)
{
    std::lock_guard<std::mutex> lk(io->mu);
    io->slot[s].busy = false;
    io->cv.notify_all();
}

// The slot's device bytes, once the solver's stream has produced them, to
// its host half on the io stream.
int copy_out(nova_io *io, int s, size_t bytes)
{
    IoSlot &sl = io->slot[s];
    NOVA_HIP(hipEventRecord(io->ev_in, io->ctx->stream));
    NOVA_HIP(hipStreamWaitEvent(io->st, io->ev_in, 0));
    NOVA_HIP(hipMemcpyAsync(sl.host, sl.dev, bytes, hipMemcpyDeviceToHost, io->st));
    NOVA_HIP(hipEventRecord(sl.ev, io->st));
    return NOVA_OK;
}

int launch_ens_pack(nova_io *io, const double *F, int nc, float *out)
{
    const nova_ctx *ctx = io->ctx;
    const EnsDims d = ens_dims(ctx);
    const int ntk = (d.e[2] + kTile - 1) / kTile, nti = (d.e[0] + kTile - 1) / kTile;
    TimedScope ts(io->ctx, "ensight_pack");
    hipLaunchKernelGGL(k_ens_pack, dim3(ntk * nti, d.e[1], nc), dim3(kTile, kTileRows), 0, ctx->stream, ctx->g, d, F,
                       out);
    return check_launch("ensight_pack");
}

}  // namespace

// ---- the clock (src/Time/Time.cpp) ----

int time_create(double dt, double end, int write_every, double interval, bool timebased, nova_time **out)
{
    auto *t = new nova_time();
    t->dt = dt;
    t->end = end;
    t->write_every = write_every;
    t->timebased = timebased;
    t->interval = interval;
    *out = t;
    return NOVA_OK;
}

// Time.cpp:77-109: "dt:", "writeT:", "writeN:", "end:" lines.
int time_create_from_file(const char *file, nova_time **out)
{
    std::ifstream fin(file);
    if (!fin) return fail(NOVA_ERR_IO, std::string("cannot open ") + file);
    auto *t = new nova_time();
    std::string line;
    while (std::getline(fin, line)) {
        std::stringstream sp(line);
        std::string key;
        sp >> key;
        if (key == "dt:") sp >> t->dt;
        else if (key == "writeT:") {
            t->timebased = true;
            sp >> t->interval;
        } else if (key == "writeN:") {
            t->timebased = false;
            sp >> t->write_every;
        } else if (key == "end:") sp >> t->end;
    }
    *out = t;
    return NOVA_OK;
}

// Time.cpp:40-61 (settings::restart()): the count, then the output times;
// the clock resumes at the last one, with as many files written.
int time_restart(nova_time *t, const char *file)
{
    std::ifstream fin(file);
    if (!fin) return fail(NOVA_ERR_IO, std::string("cannot open ") + file);
    unsigned n = 0;
    if (!(fin >> n)) return fail(NOVA_ERR_IO, std::string("bad restart file ") + file);
    std::vector<double> v;
    for (unsigned i = 0; i < n; i++) {
        double x;
        if (!(fin >> x)) return fail(NOVA_ERR_IO, std::string("bad restart file ") + file);
        v.push_back(x);
    }
    t->steps.insert(t->steps.end(), v.begin(), v.end());
    if (n) t->cur = v.back();
    t->file_num = (int)t->steps.size();
    return NOVA_OK;
}

// Time.cpp:112-139 (operator++).
int time_advance(nova_time *t)
{
    if (t->timebased) {
        t->write = (t->cur + t->dt) >= (t->file_num + 1) * t->interval;
    } else {
        if (t->write_every <= 0) return fail(NOVA_ERR_ARG, "step-count output needs writeN > 0");
        t->write = (t->time_step % t->write_every == 0);
    }
    if (t->write) {
        t->file_num++;
        t->steps.push_back(t->cur);
    }
    t->log = (t->time_step % 100 == 0);
    t->cur += t->dt;
    t->time_step++;
    return NOVA_OK;
}

void time_destroy(nova_time *t) { delete t; }

void time_set_dt(nova_time *t, double dt) { t->dt = dt; }

void time_info(const nova_time *t, double dout[4], int64_t iout[5])
{
    dout[0] = t->dt;
    dout[1] = t->cur;
    dout[2] = t->end;
    dout[3] = t->interval;
    iout[0] = t->time_step;
    iout[1] = t->file_num;
    iout[2] = t->write ? 1 : 0;
    iout[3] = t->log ? 1 : 0;
    iout[4] = (int64_t)t->steps.size();
}

void time_steps(const nova_time *t, double *out, int64_t cap)
{
    const int64_t n = (int64_t)t->steps.size();
    for (int64_t i = 0; i < n && i < cap; i++) out[i] = t->steps[i];
}

// ---- io object ----

int io_create(nova_ctx *ctx, nova_io **out)
{
    NOVA_HIP(hipSetDevice(ctx->device));
    auto *io = new nova_io();
    io->ctx = ctx;
    int rc = NOVA_OK;
    if (hipStreamCreateWithFlags(&io->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&io->ev_in, hipEventDisableTiming) != hipSuccess)
        rc = fail(NOVA_ERR_HIP, "io stream or event");
    for (auto &s : io->slot)
        if (!rc && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess)
            rc = fail(NOVA_ERR_HIP, "io event");
    if (rc) {
        for (auto &sl : io->slot)
            if (sl.ev) hipEventDestroy(sl.ev);
        if (io->ev_in) hipEventDestroy(io->ev_in);
        if (io->st) hipStreamDestroy(io->st);
        delete io;
        return rc;
    }
    io->worker = std::thread(worker_main, io);
    *out = io;
    return NOVA_OK;
}

int io_wait(nova_io *io)
{
    std::unique_lock<std::mutex> lk(io->mu);
    io->cv.wait(lk, [&] { return io->pending == 0; });
    if (io->err_code) {
        const int c = io->err_code;
        io->err_code = 0;
        return fail(c, io->err);
    }
    return NOVA_OK;
}

int io_destroy(nova_io *io)
{
    const int rc = io_wait(io);
    {
        std::lock_guard<std::mutex> lk(io->mu);
        io->stop = true;
        io->cv.notify_all();
    }
    io->worker.join();
    for (auto &s : io->slot) {
        if (s.dev) hipFree(s.dev);
        if (s.host) hipHostFree(s.host);
        if (s.ev) hipEventDestroy(s.ev);
    }
    if (io->ev_in) hipEventDestroy(io->ev_in);
    if (io->st) hipStreamDestroy(io->st);
    delete io;
    return rc;
}

// Field::write (Field.hxx:403-451): nothing unless the clock writes this
// step; path/name.PPPP.FFFFF.dat with the description, part and point
// headers, then every component's N+1 x N+1 x N+1 nodes as float32, i
// fastest.
int io_write_field(nova_io *io, const nova_time *t, const char *path, const char *name, const double *F, int nc)
{
    if (!t->write) return NOVA_OK;
    const EnsDims d = ens_dims(io->ctx);
    const size_t bytes = sizeof(float) * (size_t)nc * ens_nodes(d);
    int s = -1;
    if (int rc = acquire_slot(io, bytes, s)) return rc;
    if (int rc = launch_ens_pack(io, F, nc, (float *)io->slot[s].dev)) {
        release_slot(io, s);
        return rc;
    }
    if (int rc = copy_out(io, s, bytes)) {
        release_slot(io, s);
        return rc;
    }
    IoJob job;
    job.file = std::string(path) + "/" + name + "." + num4(kProcNo) + "." + num5(t->file_num) + ".dat";
    put80(job.head, "description");
    put80(job.head, "part");
    put<int>(job.head, 1);
    put80(job.head, "point");
    job.slot = s;
    job.bytes = bytes;
    enqueue(io, std::move(job));
    return NOVA_OK;
}

// Mesh::write (Mesh.cpp:224-303), then writeCase (:306-347) and
// writeRestartInfo (:349-365).
int io_write_mesh(nova_io *io, const nova_time *t, const char *path)
{
    if (!t->write) return NOVA_OK;
    const nova_ctx *ctx = io->ctx;
    const EnsDims d = ens_dims(ctx);
    const std::string pn = num4(kProcNo);
    IoJob geo;
    geo.file = std::string(path) + "/geom.p" + pn + ".geo";
    for (const char *s : {"C Binary", "desc1", "desc2", "node id assign", "element id assign", "part"}) put80(geo.head, s);
    put<int>(geo.head, 1);
    put80(geo.head, "part desc");
    put80(geo.head, "block rectilinear");
    for (int a = 0; a < 3; a++) put<int>(geo.head, d.e[a]);
    for (int i = 0; i < d.e[0]; i++) put<float>(geo.head, (float)(ctx->desc.origin[0] + ctx->d[0] * i));
    for (int j = 0; j < d.e[1]; j++) put<float>(geo.head, (float)(ctx->desc.origin[1] + ctx->d[1] * j));
    for (int k = 0; k < d.e[2]; k++) put<float>(geo.head, (float)mesh_z_of(ctx, k + ctx->g.h));
    enqueue(io, std::move(geo));

    std::ostringstream cs;
    cs << "FORMAT" << std::endl;
    cs << "type: ensight gold" << std::endl;
    cs << "GEOMETRY" << std::endl;
    cs << "model: geom.p" << pn << ".geo" << std::endl;
    cs << "VARIABLE" << std::endl;
    cs << "vector per node: velocity U." << pn << ".*****.dat" << std::endl;
    cs << "vector per node: Uprime Uprime." << pn << ".*****.dat" << std::endl;
    cs << "vector per node: Udash Udash." << pn << ".*****.dat" << std::endl;
    cs << "scalar per node: p  p." << pn << ".*****.dat" << std::endl;
    cs << "scalar per node: T  T." << pn << ".*****.dat" << std::endl;
    cs << "vector per node: B B." << pn << ".*****.dat" << std::endl;
    cs << "vector per node: J J." << pn << ".*****.dat" << std::endl;
    cs << "TIME" << std::endl;
    cs << "time set: 1" << std::endl;
    cs << "number of steps: " << t->steps.size() << std::endl;
    cs << "filename start number: 00001" << std::endl;
    cs << "filename increment: 1" << std::endl;
    cs << "time values: ";
    for (double v : t->steps) cs << v << std::endl;
    cs << std::endl;
    IoJob cj;
    cj.file = std::string(path) + "/fd.p" + pn + ".case";
    cj.head = cs.str();
    cj.text = true;
    enqueue(io, std::move(cj));

    std::ostringstream rs;
    rs << t->steps.size() << std::endl;
    for (double v : t->steps) rs << v << std::endl;
    IoJob rj;
    rj.file = std::string(path) + "/restart.p." + pn + ".case";
    rj.head = rs.str();
    rj.text = true;
    enqueue(io, std::move(rj));
    return NOVA_OK;
}

// Field::read (Field.hxx:455-491, :494-522 with an explicit number): skip the
// 244 header bytes, read the float32 nodes back (the periodic image entries
// are skipped).  Waits for queued writes first.
int io_read_field(nova_io *io, const char *path, const char *name, int file_num, double *F, int nc)
{
    if (int rc = io_wait(io)) return rc;
    const nova_ctx *ctx = io->ctx;
    const EnsDims d = ens_dims(ctx);
    const size_t bytes = sizeof(float) * (size_t)nc * ens_nodes(d);
    const std::string file =
        std::string(path) + "/" + name + "." + num4(kProcNo) + "." + num5(file_num) + ".dat";
    FILE *f = std::fopen(file.c_str(), "rb");
    if (!f) return fail(NOVA_ERR_IO, "cannot open " + file);
    int s = -1;
    int rc = acquire_slot(io, bytes, s);
    if (!rc) {
        if (std::fseek(f, 240 + sizeof(int), SEEK_SET) != 0 || std::fread(io->slot[s].host, 1, bytes, f) != bytes)
            rc = fail(NOVA_ERR_IO, "short file " + file);
    }
    std::fclose(f);
    if (!rc) {
        const int ntk = (ctx->g.s[2] + kTile - 1) / kTile, nti = (ctx->g.s[0] + kTile - 1) / kTile;
        if (hipMemcpyAsync(io->slot[s].dev, io->slot[s].host, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(NOVA_ERR_HIP, "host-to-device copy");
        if (!rc) {
            TimedScope ts(io->ctx, "ensight_unpack");
            hipLaunchKernelGGL(k_ens_unpack, dim3(ntk * nti, ctx->g.s[1], nc), dim3(kTile, kTileRows), 0, ctx->stream,
                               ctx->g, d, (const float *)io->slot[s].dev, F);
            rc = check_launch("ensight_unpack");
        }
        // the slot's host half is reused by the next write
        if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(NOVA_ERR_HIP, "read sync");
    }
    if (s >= 0) release_slot(io, s);
    return rc;
}

// ---- FP64 checkpoints (no reference counterpart: the reference restarts
// from its float32 EnSight files) ----
//
// Layout: the 64-byte header below, the component count of each field
// (int32), then each field's components in the stored layout (k fastest,
// no periodic image), FP64, little-endian.
struct CkptHead {
    char magic[8];
    int32_t version;
    int32_t n[3];
    int32_t wall[3];
    int32_t m;
    int32_t nfields;
    int32_t pad;
    double time;
    int64_t step;
};
static_assert(sizeof(CkptHead) == 64, "checkpoint header");
constexpr char kCkptMagic[8] = {'N', 'O', 'V', 'A', 'F', 'P', '6', '4'};

int io_checkpoint_write(nova_io *io, const char *file, int nfields, const double *const *F, const int *nc,
                        double time, int64_t step)
{
    const nova_ctx *ctx = io->ctx;
    size_t bytes = 0;
    for (int f = 0; f < nfields; f++) bytes += sizeof(double) * (size_t)nc[f] * ctx->g.np;
    int s = -1;
    if (int rc = acquire_slot(io, bytes, s)) return rc;
    size_t off = 0;
    for (int f = 0; f < nfields; f++) {
        const size_t b = sizeof(double) * (size_t)nc[f] * ctx->g.np;
        if (hipMemcpyAsync((char *)io->slot[s].dev + off, F[f], b, hipMemcpyDeviceToDevice, ctx->stream) !=
            hipSuccess) {
            release_slot(io, s);
            return fail(NOVA_ERR_HIP, "checkpoint staging copy");
        }
        off += b;
    }
    if (int rc = copy_out(io, s, bytes)) {
        release_slot(io, s);
        return rc;
    }
    CkptHead h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    for (int a = 0; a < 3; a++) {
        h.n[a] = ctx->g.n[a];
        h.wall[a] = ctx->g.wall[a];
    }
    h.m = ctx->m;
    h.nfields = nfields;
    h.time = time;
    h.step = step;
    IoJob job;
    job.file = file;
    put(job.head, h);
    for (int f = 0; f < nfields; f++) put<int32_t>(job.head, nc[f]);
    job.slot = s;
    job.bytes = bytes;
    enqueue(io, std::move(job));
    return NOVA_OK;
}

int io_checkpoint_read(nova_io *io, const char *file, int nfields, double *const *F, const int *nc, double *time,
                       int64_t *step)
{
    if (int rc = io_wait(io)) return rc;
    nova_ctx *ctx = io->ctx;
    FILE *f = std::fopen(file, "rb");
    if (!f) return fail(NOVA_ERR_IO, std::string("cannot open ") + file);
    CkptHead h{};
    std::vector<int32_t> fnc(nfields > 0 ? nfields : 1);
    int rc = NOVA_OK;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1)
        rc = fail(NOVA_ERR_IO, std::string("not a checkpoint: ") + file);
    if (!rc) {
        bool same = h.m == ctx->m && h.nfields == nfields;
        for (int a = 0; a < 3; a++) same = same && h.n[a] == ctx->g.n[a] && h.wall[a] == ctx->g.wall[a];
        if (same && std::fread(fnc.data(), sizeof(int32_t), nfields, f) != (size_t)nfields) same = false;
        for (int q = 0; same && q < nfields; q++) same = fnc[q] == nc[q];
        if (!same) rc = fail(NOVA_ERR_ARG, std::string("checkpoint does not match the grid or fields: ") + file);
    }
    size_t bytes = 0;
    for (int q = 0; q < nfields; q++) bytes += sizeof(double) * (size_t)nc[q] * ctx->g.np;
    int s = -1;
    if (!rc) rc = acquire_slot(io, bytes, s);
    if (!rc && std::fread(io->slot[s].host, 1, bytes, f) != bytes) rc = fail(NOVA_ERR_IO, std::string("short checkpoint ") + file);
    std::fclose(f);
    size_t off = 0;
    for (int q = 0; !rc && q < nfields; q++) {
        const size_t b = sizeof(double) * (size_t)nc[q] * ctx->g.np;
        if (hipMemcpyAsync(F[q], (char *)io->slot[s].host + off, b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(NOVA_ERR_HIP, "checkpoint host-to-device copy");
        off += b;
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(NOVA_ERR_HIP, "checkpoint sync");
    if (s >= 0) release_slot(io, s);
    if (!rc) {
        if (time) *time = h.time;
        if (step) *step = h.step;
    }
    return rc;
}

}  // namespace nova
