// nova.hpp -- C++ drop-in mirror of the reference's host API over the C ABI.
//
// The reference apps are written against Field<T>, Mesh, Time, settings,
// the ex:: operators, poisson and tools (src/Field/Field.h:56-367,
// src/Mesh/Mesh.h:36-178, src/Time/Time.h:29-93,
// src/settings/settings.h:28-87, src/ex/*, src/poisson/poisson.h:36-115,
// src/Tools/Tools.h:8-37).  This header keeps those names, argument
// meanings and index conventions (padded node indices, k fastest, ghosts of
// width m/2 plus the duplicated periodic endpoint), while the data lives on
// the GPU in the compact layout and every operator is a gfx950 kernel
// reached through include/nova_capi.h.  A host mirror in the reference's
// padded array-of-structures layout backs operator()(i,j,k); it is filled
// from the device (ghosts included) on first host access after a device
// write, and uploaded before the next device operation after host writes.
//
// Errors from the library throw nova::Error (the reference asserts).
#pragma once

#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "nova_capi.h"

typedef double scalar;

class vector {
    scalar v_[3];

public:
    vector() : v_{0.0, 0.0, 0.0} {}
    vector(scalar a) : v_{a, a, a} {}
    vector(scalar a, scalar b, scalar c) : v_{a, b, c} {}
    scalar x() const { return v_[0]; }
    scalar y() const { return v_[1]; }
    scalar z() const { return v_[2]; }
    scalar &x() { return v_[0]; }
    scalar &y() { return v_[1]; }
    scalar &z() { return v_[2]; }
    scalar operator[](int i) const { return v_[i]; }
    scalar &operator[](int i) { return v_[i]; }
};

// src/Types/tensor/tensor.h:25: nine components, row-major (xx, xy, xz, yx, ...)
class tensor {
    scalar v_[9];

public:
    tensor() : v_{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
    tensor(scalar a) : v_{a, a, a, a, a, a, a, a, a} {}
    scalar operator[](int i) const { return v_[i]; }
    scalar &operator[](int i) { return v_[i]; }
    scalar xx() const { return v_[0]; }
    scalar xy() const { return v_[1]; }
    scalar xz() const { return v_[2]; }
    scalar yx() const { return v_[3]; }
    scalar yy() const { return v_[4]; }
    scalar yz() const { return v_[5]; }
    scalar zx() const { return v_[6]; }
    scalar zy() const { return v_[7]; }
    scalar zz() const { return v_[8]; }
};

template <class T> struct pTraits;
template <> struct pTraits<scalar> { static constexpr int nComp = 1; };
template <> struct pTraits<vector> { static constexpr int nComp = 3; };
template <> struct pTraits<tensor> { static constexpr int nComp = 9; };

namespace nova {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what)
{
    if (rc != NOVA_OK) throw Error(std::string(what) + ": " + nova_last_error());
}

}  // namespace nova

// ---- settings (src/settings/settings.cpp:21-77) ----
class settings {
public:
    static int &m()
    {
        static int m_ = 6;
        return m_;
    }
    static std::string &zoneName()
    {
        static std::string z = ".";
        return z;
    }
    static bool &restart()
    {
        static bool r = false;
        return r;
    }
    // central-difference weights of orders 2, 4, 6, 8
    static const std::array<scalar, 4> &coef()
    {
        static const std::array<std::array<scalar, 4>, 4> c = {
            std::array<scalar, 4>{{1.0, 0.0, 0.0, 0.0}}, std::array<scalar, 4>{{4.0 / 3.0, -1.0 / 3.0, 0.0, 0.0}},
            std::array<scalar, 4>{{6.0 / 4.0, -12.0 / 20.0, 1.0 / 10.0, 0.0}},
            std::array<scalar, 4>{{8.0 / 5.0, -4.0 / 5.0, 24.0 / 105.0, -8.0 / 280.0}}};
        return c[m() / 2 - 1];
    }
    // RK4 stage table and stage times
    static const std::array<std::array<scalar, 4>, 4> &RK4a()
    {
        static const std::array<std::array<scalar, 4>, 4> a = {
            std::array<scalar, 4>{{0.5, 0, 0, 0}}, std::array<scalar, 4>{{0, 0.5, 0, 0}},
            std::array<scalar, 4>{{0, 0, 1.0, 0}},
            std::array<scalar, 4>{{1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}}};
        return a;
    }
    static const std::array<scalar, 4> &RK4c()
    {
        static const std::array<scalar, 4> c = {0.5, 0.5, 1.0, 1.0};
        return c;
    }
    // command line: -restart, -zone <dir>, -m <even order>
    static void process(int argc, char *argv[])
    {
        for (int i = 1; i < argc; i++) {
            const std::string a = argv[i];
            if (a == "-restart") restart() = true;
            if (a == "-zone" && i + 1 < argc) zoneName() = argv[i + 1];
            if (a == "-m" && i + 1 < argc) {
                m() = std::atoi(argv[i + 1]);
                if (m() % 2 || m() < 2 || m() > 8) throw nova::Error("-m must be 2, 4, 6 or 8");
            }
        }
    }
};

// ---- Time (src/Time/Time.cpp:22-139): the library's clock ----
// Time(dt, endTime, writeEvery): output every writeEvery steps from the
// first; Time(dt, endTime, writeInterval): output when the time reaches the
// next multiple of the interval; Time(file): "dt:", "writeT:", "writeN:",
// "end:" lines.  With settings::restart() the clock resumes from
// ./data/restart.pPPPP.case, the name Time.cpp:44 reads (Mesh::write names
// its restart file restart.p.PPPP.case).
class Time {
    nova_time *t_ = nullptr;
    scalar dt_ = 0.0;  // dt() is writable; pushed to the clock before each step

    void restartIfAsked()
    {
        if (settings::restart()) nova::check(nova_time_restart(t_, "./data/restart.p0000.case"), "Time restart");
    }
    void info(double d[4], int64_t i[5]) const { nova::check(nova_time_info(t_, d, i), "nova_time_info"); }

public:
    Time(scalar dt, scalar endTime, int writeEvery) : dt_(dt)
    {
        nova::check(nova_time_create(dt, endTime, writeEvery, &t_), "nova_time_create");
        restartIfAsked();
    }
    Time(scalar dt, scalar endTime, scalar writeInterval) : dt_(dt)
    {
        nova::check(nova_time_create_interval(dt, endTime, writeInterval, &t_), "nova_time_create_interval");
        restartIfAsked();
    }
    explicit Time(const std::string &fileName)
    {
        nova::check(nova_time_create_from_file(fileName.c_str(), &t_), "nova_time_create_from_file");
        double d[4];
        int64_t i[5];
        info(d, i);
        dt_ = d[0];
    }
    Time(const Time &) = delete;
    Time &operator=(const Time &) = delete;
    ~Time() { nova_time_destroy(t_); }

    scalar &dt() { return dt_; }
    bool run() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return d[1] < d[2];
    }
    bool write() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return i[2] != 0;
    }
    bool writelog() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return i[3] != 0;
    }
    scalar curTime() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return d[1];
    }
    int timeStep() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return (int)i[0];
    }
    int fileNum() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        return (int)i[1];
    }
    std::vector<scalar> steps() const
    {
        double d[4];
        int64_t i[5];
        info(d, i);
        std::vector<scalar> v((size_t)i[4]);
        nova::check(nova_time_steps(t_, v.data(), i[4]), "nova_time_steps");
        return v;
    }
    void operator++(int)
    {
        nova::check(nova_time_set_dt(t_, dt_), "nova_time_set_dt");
        nova::check(nova_time_advance(t_), "nova_time_advance");
    }
    const nova_time *handle() const { return t_; }
};

// single-rank reductions (src/parallelCom/parallelCom.h:244-282)
template <class T> struct plusOp { T operator()(T a, T b) const { return a + b; } };
template <class T> struct maxOp { T operator()(T a, T b) const { return a > b ? a : b; } };
template <class T> struct minOp { T operator()(T a, T b) const { return a < b ? a : b; } };
template <class T, class Op> inline void reduce(T &, Op) {}
template <class T, class Op> inline void all_reduce(T &, Op) {}

// ---- Mesh (src/Mesh/Mesh.cpp:24-221): grid + device context ----
class Mesh {
    struct Ctx {
        nova_ctx *ctx = nullptr;
        nova_poisson *ps = nullptr;
        nova_io *io = nullptr;
        std::map<int, std::vector<double *>> pool;  // free device buffers by component count
        // the reference's wall fills per walled direction, from the BC files:
        // 0 none, 1 implied by a scalar field's file, 2 given by a vector's
        int wallSet[3] = {0, 0, 0};
        int wallKind[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        ~Ctx()
        {
            if (io) nova_io_destroy(io);  // waits for queued writes
            for (auto &kv : pool)
                for (double *p : kv.second) nova_free(ctx, p);
            if (ps) nova_poisson_destroy(ps);
            if (ctx) nova_ctx_destroy(ctx);
        }
    };
    std::shared_ptr<Ctx> c_;
    nova_grid_desc d_{};
    std::string zone_ = ".";  // where the <name>BC files are read
    int n_[3] = {0, 0, 0};  // padded extents
    int64_t np_ = 0;
    Time *time_ = nullptr;

    void open(int device)
    {
        c_ = std::make_shared<Ctx>();
        nova::check(nova_ctx_create(&d_, device, &c_->ctx), "nova_ctx_create");
        int64_t info[8];
        nova::check(nova_ctx_info(c_->ctx, info), "nova_ctx_info");
        np_ = info[3];
        n_[0] = (int)info[4];
        n_[1] = (int)info[5];
        n_[2] = (int)info[6];
    }

public:
    // mesh file keys "origin:", "N:", "L:", "hypz:" (Mesh.cpp:38-62); a
    // non-periodic north/south entry in <zone>/UBC makes y walled.
    Mesh(const std::string &fileName, Time &runTime, int device = 0) : time_(&runTime)
    {
        std::ifstream fin(fileName);
        if (!fin) throw nova::Error("cannot open mesh file " + fileName);
        std::string line;
        while (std::getline(fin, line)) {
            std::stringstream sp(line);
            std::string key;
            sp >> key;
            if (key == "origin:") sp >> d_.origin[0] >> d_.origin[1] >> d_.origin[2];
            else if (key == "N:") sp >> d_.n[0] >> d_.n[1] >> d_.n[2];
            else if (key == "L:") sp >> d_.L[0] >> d_.L[1] >> d_.L[2];
            else if (key == "hypz:") { d_.hyp = 1; sp >> d_.s1 >> d_.s2; }
        }
        d_.m = settings::m();
        const std::string zone = fileName.substr(0, fileName.find_last_of('/') == std::string::npos
                                                        ? 0 : fileName.find_last_of('/'));
        // a direction whose UBC entries are not periodic is walled
        // (src/BC/BCenums.cpp:3-5 names: east/west = x, north/south = y,
        // top/bottom = z)
        std::ifstream ubc((zone.empty() ? std::string(".") : zone) + "/UBC");
        std::string dir, type;
        while (ubc >> dir >> type) {
            const int a = (dir == "east" || dir == "west") ? 0 : (dir == "north" || dir == "south") ? 1 : 2;
            if (type != "periodic") d_.bc[a] = NOVA_BC_WALL;
            std::getline(ubc, line);
        }
        zone_ = zone.empty() ? std::string(".") : zone;
        open(device);
    }
    // zone: where the fields' <name>BC files are read
    Mesh(const std::array<int, 3> &N, const std::array<scalar, 3> &L, Time &runTime,
         const std::array<scalar, 3> &origin = {0.0, 0.0, 0.0}, bool wallY = false, int device = 0,
         const std::string &zone = ".")
        : zone_(zone), time_(&runTime)
    {
        for (int a = 0; a < 3; a++) {
            d_.n[a] = N[a];
            d_.L[a] = L[a];
            d_.origin[a] = origin[a];
        }
        d_.m = settings::m();
        d_.bc[1] = wallY ? NOVA_BC_WALL : NOVA_BC_PERIODIC;
        open(device);
    }
    int ni() const { return n_[0]; }
    int nj() const { return n_[1]; }
    int nk() const { return n_[2]; }
    bool walled(int a) const { return d_.bc[a] == NOVA_BC_WALL; }
    const std::string &zone() const { return zone_; }
    scalar dx() const { return d_.L[0] / d_.n[0]; }
    scalar dy() const { return d_.L[1] / d_.n[1]; }
    scalar dz() const { return d_.L[2] / d_.n[2]; }
    scalar lx() const { return d_.L[0]; }
    scalar ly() const { return d_.L[1]; }
    scalar lz() const { return d_.L[2]; }
    vector origin() const { return vector(d_.origin[0], d_.origin[1], d_.origin[2]); }
    Time &time() const { return *time_; }
    nova_ctx *ctx() const { return c_->ctx; }
    int64_t points() const { return np_; }
    nova_poisson *poissonHandle() const
    {
        if (!c_->ps) {
            nova::check(nova_poisson_create(c_->ctx, &c_->ps), "nova_poisson_create");
            if (walled(1) && c_->wallSet[1])  // the reference's FFT solve across the walls
                nova::check(nova_poisson_set_wall_solve(c_->ps, NOVA_WALL_SOLVE_REFERENCE), "poisson walls");
        }
        return c_->ps;
    }
    // The reference's fills on the walls normal to a (dirichletBC /
    // neumannBC with value 0, per vector component), installed by the BC
    // files: every vector field of the mesh shares them, since the device
    // reflects a component the same way in every stencil.  A scalar field's
    // file (neumann 0) installs the fills with the wall-normal component
    // odd until a vector field's file names the components.
    void setWallBC(int a, const std::array<int, 3> &kind, bool fromVector, const std::string &file)
    {
        if (fromVector && c_->wallSet[a] == 2) {
            for (int c = 0; c < 3; c++)
                if (c_->wallKind[a][c] != kind[c])
                    throw nova::Error(file + ": the vector fields of one mesh must share their wall conditions");
            return;
        }
        if (!fromVector && c_->wallSet[a]) return;
        nova::check(nova_ctx_set_wall_bc(c_->ctx, a, kind.data()), "nova_ctx_set_wall_bc");
        for (int c = 0; c < 3; c++) c_->wallKind[a][c] = kind[c];
        c_->wallSet[a] = fromVector ? 2 : 1;
        if (c_->ps && a == 1)
            nova::check(nova_poisson_set_wall_solve(c_->ps, NOVA_WALL_SOLVE_REFERENCE), "poisson walls");
    }
    // pool-owned device buffers (the reference's reuseTmp, reuseTmp.hxx:19-50)
    double *take(int nc) const
    {
        auto &v = c_->pool[nc];
        if (!v.empty()) {
            double *p = v.back();
            v.pop_back();
            return p;
        }
        double *p = nullptr;
        nova::check(nova_alloc(c_->ctx, nc, &p), "nova_alloc");
        return p;
    }
    void give(int nc, double *p) const { c_->pool[nc].push_back(p); }
    // output (asynchronous; files land in the order written)
    nova_io *ioHandle() const
    {
        if (!c_->io) nova::check(nova_io_create(c_->ctx, &c_->io), "nova_io_create");
        return c_->io;
    }
    // Mesh::write (Mesh.cpp:224-303) with writeCase and writeRestartInfo
    void write(const std::string &path) const
    {
        nova::check(nova_io_write_mesh(ioHandle(), time_->handle(), path.c_str()), "Mesh::write");
    }
    // blocks until every queued file is written
    void flushOutput() const
    {
        if (c_->io) nova::check(nova_io_wait(c_->io), "nova_io_wait");
    }
    std::shared_ptr<Ctx> keep() const { return c_; }
};

template <class T>
class Field {
    static constexpr int NC = pTraits<T>::nComp;
    // the host mirror is read and written as NC consecutive doubles per node
    static_assert(sizeof(T) == NC * sizeof(double), "node types are NC packed doubles");
    Mesh m_;
    std::string name_;
    double *d_ = nullptr;                 // device, compact SoA
    // host mirror, padded AoS, in page-locked memory (allocated at the first
    // host access): the transfers to and from it run at the link's rate,
    // about three times the rate through pageable memory
    struct HostFree {
        void operator()(T *p) const { nova_host_free(p); }
    };
    static_assert(std::is_trivially_destructible<T>::value, "node types hold plain doubles");
    mutable std::unique_ptr<T[], HostFree> h_;
    mutable size_t hn_ = 0;
    mutable bool hostValid_ = false;      // mirror equals device
    mutable bool hostDirty_ = false;      // mirror newer than device

    void syncHost() const
    {
        if (hostValid_) return;
        if (!h_) {
            const size_t n = (size_t)m_.ni() * m_.nj() * m_.nk();
            void *p = nullptr;
            nova::check(nova_host_alloc(n * sizeof(T), &p), "nova_host_alloc");
            h_.reset(static_cast<T *>(p));
            std::uninitialized_default_construct_n(h_.get(), n);
            hn_ = n;
        }
        const size_t np = hn_ * NC;
        double *dpad = padBuffer();
        nova::check(nova_pack_padded(m_.ctx(), d_, NC, dpad), "nova_pack_padded");
        // the node types are NC consecutive doubles: copy straight into the mirror
        nova::check(nova_memcpy_d2h(m_.ctx(), (double *)h_.get(), dpad, np), "nova_memcpy_d2h");
        releasePad(dpad);
        hostValid_ = true;
    }
    double *padBuffer() const
    {
        // padded extent in units of the compact field: round up
        const int64_t padded = (int64_t)m_.ni() * m_.nj() * m_.nk() * NC;
        const int nc = (int)((padded + m_.points() - 1) / m_.points());
        return m_.take(nc);
    }
    void releasePad(double *p) const
    {
        const int64_t padded = (int64_t)m_.ni() * m_.nj() * m_.nk() * NC;
        m_.give((int)((padded + m_.points() - 1) / m_.points()), p);
    }

public:
    // device-side synchronisation (host writes -> device)
    void syncDevice() const
    {
        if (!hostDirty_) return;
        const size_t np = hn_ * NC;
        double *dpad = padBuffer();
        nova::check(nova_memcpy_h2d(m_.ctx(), dpad, (const double *)h_.get(), np), "nova_memcpy_h2d");
        nova::check(nova_unpack_padded(m_.ctx(), dpad, NC, d_), "nova_unpack_padded");
        releasePad(dpad);
        hostDirty_ = false;
    }
    void deviceWritten() { hostValid_ = false; hostDirty_ = false; }

    Field(Mesh &m, const T &init, std::string name = "") : m_(m), name_(std::move(name))
    {
        d_ = m_.take(NC);
        *this = init;
        if (!name_.empty()) setBC(name_ + "BC");
    }
    // Field::setBC (src/Field/Field.hxx:519-547): the "dir type val comp"
    // lines of <zone>/<file> (a missing file leaves the field as it is).
    // Periodic directions take "periodic".  On a walled direction every
    // component takes "dirichlet 0" or "neumann 0" (dirichletBC.hxx:90-125,
    // neumannBC.hxx:100-135): the stencils reflect the component oddly or
    // evenly about the wall node and correctBoundaryConditions() overwrites
    // the wall node as those fills do; scalar fields take "neumann 0".
    // Anything else (non-zero values, a periodic entry on a walled direction
    // or the reverse, a vector missing a component's entry) cannot be
    // represented and throws.
    void setBC(const std::string &file)
    {
        std::ifstream fin(m_.zone() + "/" + file);
        if (!fin.is_open()) return;
        std::string line;
        static const std::map<std::string, int> axis = {{"east", 0}, {"west", 0}, {"north", 1},
                                                        {"south", 1}, {"top", 2}, {"bottom", 2}};
        int kind[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        while (std::getline(fin, line)) {
            std::stringstream sp(line);
            std::string dir, type;
            scalar val = 0.0;
            int comp = 0;
            if (!(sp >> dir >> type >> val >> comp)) continue;
            const auto it = axis.find(dir);
            if (it == axis.end()) throw nova::Error(file + ": unknown direction " + dir);
            const int a = it->second;
            if (comp < 0 || comp >= NC) throw nova::Error(file + ": component " + std::to_string(comp) + " out of range");
            const bool ok = m_.walled(a) ? ((type == "neumann" || (NC == 3 && type == "dirichlet")) && val == 0.0)
                                         : type == "periodic";
            if (!ok)
                throw nova::Error(file + ": '" + line + "' is not a condition the device applies on a " +
                                  (m_.walled(a) ? "walled direction (dirichlet 0 or neumann 0 per vector component, "
                                                  "neumann 0 for scalars)"
                                                : "periodic direction (periodic)"));
            if (m_.walled(a)) {
                const int k = type == "dirichlet" ? NOVA_WALL_DIRICHLET0 : NOVA_WALL_NEUMANN0;
                if (kind[a][comp] && kind[a][comp] != k)
                    throw nova::Error(file + ": the two walls of a direction must agree (" + line + ")");
                kind[a][comp] = k;
            }
        }
        for (int a = 0; a < 3; a++) {
            if (!m_.walled(a) || !kind[a][0]) continue;
            if (NC == 3) {
                if (!kind[a][1] || !kind[a][2]) throw nova::Error(file + ": a wall entry for every component is needed");
                m_.setWallBC(a, {kind[a][0], kind[a][1], kind[a][2]}, true, file);
            } else {
                std::array<int, 3> k = {NOVA_WALL_NEUMANN0, NOVA_WALL_NEUMANN0, NOVA_WALL_NEUMANN0};
                k[a] = NOVA_WALL_DIRICHLET0;
                m_.setWallBC(a, k, false, file);
            }
        }
    }
    Field(Mesh &m) : Field(m, T(0.0)) {}
    // an operator's result: storage only, every stored node is written by
    // the operator before anything reads it (the reference zero-fills its
    // pool temporaries, reuseTmp.hxx:47, which every operator then overwrites)
    struct Uninit {};
    Field(Mesh &m, Uninit) : m_(m) { d_ = m_.take(NC); }
    Field(const Field &f) : m_(f.m_), name_(f.name_)
    {
        d_ = m_.take(NC);
        *this = f;
    }
    ~Field()
    {
        if (d_) m_.give(NC, d_);
    }

    int index(int i, int j, int k) const { return k + j * nk() + i * (nk() * nj()); }
    T &operator()(int i, int j, int k) const
    {
        syncHost();
        hostDirty_ = true;
        return h_[(size_t)index(i, j, k)];
    }
    T &operator()(int i) const
    {
        syncHost();
        hostDirty_ = true;
        return h_[(size_t)i];
    }
    const Mesh &mesh() const { return m_; }
    Mesh &mesh() { return m_; }
    int ni() const { return m_.ni(); }
    int nj() const { return m_.nj(); }
    int nk() const { return m_.nk(); }
    std::string &name() { return name_; }
    const double *dev() const
    {
        syncDevice();
        return d_;
    }
    double *devw()
    {
        syncDevice();
        hostValid_ = false;
        return d_;
    }
    // ghosts are never stored on the device: every kernel applies the BC
    // on load, and the host mirror is filled with them on transfer; with the
    // reference's wall fills (installed by BC files) the wall nodes are
    // overwritten here, as the fills do
    void correctBoundaryConditions()
    {
        if (hostDirty_) syncDevice();
        hostValid_ = false;
        if (NC == 1 || NC == 3) nova::check(nova_wall_fill(m_.ctx(), d_, NC), "correctBoundaryConditions");
    }

    // Field::write / read (Field.hxx:403-522): EnSight Gold float32 files
    void write(const std::string &path, const std::string &name) const
    {
        nova::check(nova_io_write_field(m_.ioHandle(), m_.time().handle(), path.c_str(), name.c_str(), dev(), NC),
                    "Field::write");
    }
    void read(const std::string &path, const std::string &name) { read(path, name, m_.time().fileNum()); }
    void read(const std::string &path, const std::string &name, int num)
    {
        nova::check(nova_io_read_field(m_.ioHandle(), path.c_str(), name.c_str(), num, devw(), NC), "Field::read");
    }

    void operator=(const Field &f)
    {
        if (&f == this) return;
        nova::check(nova_axpby(m_.ctx(), (size_t)NC * m_.points(), 1.0, f.dev(), 0.0, nullptr, devw()), "copy");
    }
    void operator=(const std::shared_ptr<Field> &f) { *this = *f; }
    // an expression's result that nothing else holds (U = Ustar - dt*grad(p)):
    // its device buffer is taken over instead of copied, and the temporary
    // returns this field's old buffer to the pool.  The library keeps no
    // field's device address between calls, so the exchange is not visible.
    void operator=(std::shared_ptr<Field> &&f)
    {
        if (f && f.use_count() == 1 && f.get() != this && !f->hostDirty_ && f->m_.ctx() == m_.ctx()) {
            std::swap(d_, f->d_);
            hostValid_ = false;
            hostDirty_ = false;
            f->hostValid_ = false;
            return;
        }
        *this = *f;
    }
    void operator=(const T &v)
    {
        // constant fill on the device
        double cv[NC];
        for (int c = 0; c < NC; c++) cv[c] = comp(v, c);
        nova::check(nova_field_op(m_.ctx(), NOVA_OP_FILL, nullptr, NC, nullptr, 0, cv, d_), "fill");
        hostValid_ = false;
        hostDirty_ = false;
    }
    void operator+=(const Field &f) { axpby(1.0, *this, 1.0, &f); }
    void operator+=(const std::shared_ptr<Field> &f) { axpby(1.0, *this, 1.0, f.get()); }
    void operator-=(const Field &f) { axpby(1.0, *this, -1.0, &f); }
    void operator-=(const std::shared_ptr<Field> &f) { axpby(1.0, *this, -1.0, f.get()); }
    void operator*=(const scalar s) { axpby(s, *this, 0.0, nullptr); }
    void operator/=(const scalar s) { axpby(1.0 / s, *this, 0.0, nullptr); }

    static scalar comp(const T &v, int c);
    void axpby(scalar a, const Field &x, scalar b, const Field *y)
    {
        const double *xp = x.dev();
        const double *yp = y ? y->dev() : nullptr;
        nova::check(nova_axpby(m_.ctx(), (size_t)NC * m_.points(), a, xp, b, yp, devw()), "axpby");
    }
};

template <> inline scalar Field<scalar>::comp(const scalar &v, int) { return v; }
template <> inline scalar Field<vector>::comp(const vector &v, int c) { return v[c]; }
template <> inline scalar Field<tensor>::comp(const tensor &v, int c) { return v[c]; }

template <class T> using FieldPtr = std::shared_ptr<Field<T>>;

namespace nova {
template <class T>
FieldPtr<T> tmp(const Mesh &m)
{
    return std::make_shared<Field<T>>(const_cast<Mesh &>(m), typename Field<T>::Uninit{});
}
template <class T>
FieldPtr<T> lincomb(scalar a, const Field<T> &x, scalar b, const Field<T> *y)
{
    auto r = tmp<T>(x.mesh());
    r->axpby(a, x, b, y);
    return r;
}
}  // namespace nova

// ---- Field algebra returning pool temporaries (Field.h:160-367) ----
template <class T> FieldPtr<T> operator+(const Field<T> &a, const Field<T> &b) { return nova::lincomb(1.0, a, 1.0, &b); }
template <class T> FieldPtr<T> operator+(const FieldPtr<T> &a, const FieldPtr<T> &b) { return *a + *b; }
template <class T> FieldPtr<T> operator+(const Field<T> &a, const FieldPtr<T> &b) { return a + *b; }
template <class T> FieldPtr<T> operator+(const FieldPtr<T> &a, const Field<T> &b) { return *a + b; }
template <class T> FieldPtr<T> operator-(const Field<T> &a, const Field<T> &b) { return nova::lincomb(1.0, a, -1.0, &b); }
template <class T> FieldPtr<T> operator-(const FieldPtr<T> &a, const FieldPtr<T> &b) { return *a - *b; }
template <class T> FieldPtr<T> operator-(const Field<T> &a, const FieldPtr<T> &b) { return a - *b; }
template <class T> FieldPtr<T> operator-(const FieldPtr<T> &a, const Field<T> &b) { return *a - b; }
template <class T> FieldPtr<T> operator-(const FieldPtr<T> &a) { return nova::lincomb<T>(-1.0, *a, 0.0, nullptr); }
template <class T> FieldPtr<T> operator*(scalar s, const Field<T> &a) { return nova::lincomb<T>(s, a, 0.0, nullptr); }
template <class T> FieldPtr<T> operator*(scalar s, const FieldPtr<T> &a) { return s * *a; }
template <class T> FieldPtr<T> operator/(const FieldPtr<T> &a, scalar s) { return nova::lincomb<T>(1.0 / s, *a, 0.0, nullptr); }

// pointwise products (Field.h:227-363), one device pass each (nova_field_op)
namespace nova {
template <class R, class A, class B>
FieldPtr<R> fieldop(int op, const Field<A> &a, const Field<B> *b, const double *cval = nullptr)
{
    auto r = tmp<R>(a.mesh());
    check(nova_field_op(a.mesh().ctx(), op, a.dev(), pTraits<A>::nComp, b ? b->dev() : nullptr,
                        b ? pTraits<B>::nComp : 0, cval, r->devw()),
          "field operator");
    return r;
}
}  // namespace nova
// inner product (operator&, Field.h:340-363): vector & vector, scalar & scalar
template <class T> FieldPtr<scalar> operator&(const Field<T> &a, const Field<T> &b)
{
    return nova::fieldop<scalar>(NOVA_OP_DOT, a, &b);
}
template <class T> FieldPtr<scalar> operator&(const Field<T> &a, const FieldPtr<T> &b) { return a & *b; }
template <class T> FieldPtr<scalar> operator&(const FieldPtr<T> &a, const FieldPtr<T> &b) { return *a & *b; }
template <class T> FieldPtr<scalar> operator&(const FieldPtr<T> &a, const Field<T> &b) { return *a & b; }
// cross product (operator^, Field.cpp:117-150)
inline FieldPtr<vector> operator^(const Field<vector> &a, const Field<vector> &b)
{
    return nova::fieldop<vector>(NOVA_OP_CROSS, a, &b);
}
// scalar field times a field (operator*, Field.h:242-293: outerProductType<scalar, T> = T)
template <class T> FieldPtr<T> operator*(const Field<scalar> &s, const Field<T> &f)
{
    auto r = nova::tmp<T>(f.mesh());
    nova::check(nova_field_op(f.mesh().ctx(), NOVA_OP_SCALE, s.dev(), 1, f.dev(), pTraits<T>::nComp, nullptr,
                              r->devw()),
                "operator*");
    return r;
}
inline FieldPtr<vector> operator*(const Field<vector> &f, const Field<scalar> &s) { return s * f; }
template <class T> FieldPtr<T> operator*(const FieldPtr<scalar> &s, const Field<T> &f) { return *s * f; }
template <class T> FieldPtr<T> operator*(const Field<scalar> &s, const FieldPtr<T> &f) { return s * *f; }
template <class T> FieldPtr<T> operator*(const FieldPtr<scalar> &s, const FieldPtr<T> &f) { return *s * *f; }
// vector times vector: the outer product, a tensor field
inline FieldPtr<tensor> operator*(const Field<vector> &a, const Field<vector> &b)
{
    return nova::fieldop<tensor>(NOVA_OP_OUTER, a, &b);
}
// a field over a scalar field (operator/, Field.h:266-293)
template <class T> FieldPtr<T> operator/(const Field<T> &f, const Field<scalar> &s)
{
    return nova::fieldop<T>(NOVA_OP_DIVIDE, f, &s);
}
template <class T> FieldPtr<T> operator/(const FieldPtr<T> &f, const Field<scalar> &s) { return *f / s; }
template <class T> FieldPtr<T> operator/(const Field<T> &f, const FieldPtr<scalar> &s) { return f / *s; }
template <class T> FieldPtr<T> operator/(const FieldPtr<T> &f, const FieldPtr<scalar> &s) { return *f / *s; }
// tensor double inner product (operator&&, Field.h:323), as tensor.cpp:237-251 writes it
inline FieldPtr<scalar> operator&&(const FieldPtr<tensor> &a, const FieldPtr<tensor> &b)
{
    return nova::fieldop<scalar>(NOVA_OP_DDOT, *a, b.get());
}
// sqrt (Field.cpp:88-113)
inline FieldPtr<scalar> sqrt(const Field<scalar> &f) { return nova::fieldop<scalar, scalar, scalar>(NOVA_OP_SQRT, f, nullptr); }
inline FieldPtr<scalar> sqrt(const FieldPtr<scalar> &f) { return sqrt(*f); }
// a field minus a constant (operator-, Field.h:189-201)
template <class T> FieldPtr<T> operator-(const Field<T> &f, const T &v)
{
    double c[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < pTraits<T>::nComp; q++) c[q] = -Field<T>::comp(v, q);
    return nova::fieldop<T, T, T>(NOVA_OP_ADD_CONST, f, nullptr, c);
}
template <class T> FieldPtr<T> operator-(const FieldPtr<T> &f, const T &v) { return *f - v; }

// ---- explicit operators (src/ex) ----
namespace ex {
// ex::div(U, phi), src/ex/div/div.hxx:84-172
template <class T>
FieldPtr<T> div(const Field<vector> &U, const Field<T> &phi)
{
    auto r = nova::tmp<T>(U.mesh());
    nova::check(nova_div_adv(U.mesh().ctx(), U.dev(), phi.dev(), pTraits<T>::nComp, r->devw()), "ex::div");
    return r;
}
template <class T> FieldPtr<T> div(const FieldPtr<vector> &U, const FieldPtr<T> &phi) { return div(*U, *phi); }
template <class T> FieldPtr<T> div(const FieldPtr<vector> &U, const Field<T> &phi) { return div(*U, phi); }
template <class T> FieldPtr<T> div(const Field<vector> &U, const FieldPtr<T> &phi) { return div(U, *phi); }
// ex::div(U), src/ex/div/div.hxx:22-81
inline FieldPtr<scalar> div(const Field<vector> &U)
{
    auto r = nova::tmp<scalar>(U.mesh());
    nova::check(nova_div(U.mesh().ctx(), U.dev(), r->devw()), "ex::div");
    return r;
}
inline FieldPtr<scalar> div(const FieldPtr<vector> &U) { return div(*U); }
// ex::grad(p), src/ex/grad/grad.cpp:25-62
inline FieldPtr<vector> grad(const Field<scalar> &p)
{
    auto r = nova::tmp<vector>(p.mesh());
    nova::check(nova_grad(p.mesh().ctx(), p.dev(), r->devw()), "ex::grad");
    return r;
}
inline FieldPtr<vector> grad(const FieldPtr<scalar> &p) { return grad(*p); }
// ex::laplacian(nu, f), src/ex/laplacian/laplacian.hxx:130-191
template <class T>
FieldPtr<T> laplacian(const scalar &nu, const Field<T> &f)
{
    auto r = nova::tmp<T>(f.mesh());
    nova::check(nova_laplacian(f.mesh().ctx(), nu, f.dev(), pTraits<T>::nComp, r->devw()), "ex::laplacian");
    return r;
}
template <class T>
FieldPtr<T> laplacian(const scalar &nu, const FieldPtr<T> &f) { return laplacian(nu, *f); }
// ex::laplacian(Field<scalar> nu, f), src/ex/laplacian/laplacian.hxx:64-127
template <class T>
FieldPtr<T> laplacian(const Field<scalar> &nu, const Field<T> &f)
{
    auto r = nova::tmp<T>(f.mesh());
    nova::check(nova_laplacian_var(f.mesh().ctx(), nu.dev(), f.dev(), pTraits<T>::nComp, r->devw()),
                "ex::laplacian");
    return r;
}
// ex::grad(Field<vector>) -> Field<tensor>, src/ex/grad/grad.cpp:65-102
inline FieldPtr<tensor> grad(const Field<vector> &U)
{
    auto r = nova::tmp<tensor>(U.mesh());
    nova::check(nova_grad_vec_literal(U.mesh().ctx(), U.dev(), r->devw()), "ex::grad");
    return r;
}
inline FieldPtr<tensor> grad(const FieldPtr<vector> &U) { return grad(*U); }
// ex::curl, src/ex/curl/curl.cpp:25-57
inline FieldPtr<vector> curl(const Field<vector> &U)
{
    auto r = nova::tmp<vector>(U.mesh());
    nova::check(nova_curl_literal(U.mesh().ctx(), U.dev(), r->devw()), "ex::curl");
    return r;
}
}  // namespace ex

// ---- poisson (src/poisson/poisson.cpp:6-265) ----
class poisson {
    FieldPtr<scalar> p_;
    FieldPtr<scalar> rhs_;

public:
    explicit poisson(FieldPtr<scalar> p) : p_(std::move(p)) { p_->mesh().poissonHandle(); }
    void rhs(const FieldPtr<scalar> &f) { rhs_ = f; }
    void solve()
    {
        if (!rhs_) throw nova::Error("poisson::solve without rhs");
        nova::check(nova_poisson_solve(p_->mesh().poissonHandle(), rhs_->dev(), p_->devw()), "poisson::solve");
    }
};

// ---- implicit matrix (src/im/fdMatrix/fdMatrix.h, src/im/laplacian/laplacian.hxx) ----
namespace im {
template <class T>
class fdMatrix {
    static_assert(std::is_same<T, scalar>::value, "fdMatrix solves scalar fields, as the reference's does");
    nova_fdmatrix *m_ = nullptr;
    FieldPtr<scalar> fld_;  // the transported field, solved in place

public:
    fdMatrix(nova_fdmatrix *m, FieldPtr<scalar> f) : m_(m), fld_(std::move(f)) {}
    fdMatrix(const fdMatrix &) = delete;
    fdMatrix &operator=(const fdMatrix &) = delete;
    ~fdMatrix() { nova_fd_destroy(m_); }
    nova_fdmatrix *handle() const { return m_; }
    // fdMatrix.hxx:59-209
    void solve(scalar tol, int maxIts)
    {
        int its = 0;
        double res = 0.0;
        nova::check(nova_fd_solve(m_, fld_->devw(), tol, maxIts, &its, &res), "fdMatrix::solve");
        if (fld_->mesh().time().writelog()) {
            std::cout << "Poisson Equation converged in " << its << " iterations." << std::endl;
            std::cout << "Residual: " << res << "." << std::endl;
        }
    }
    void relax(const scalar &urf) { nova::check(nova_fd_relax(m_, fld_->dev(), urf), "fdMatrix::relax"); }
    T residual()
    {
        double r = 0.0;
        nova::check(nova_fd_residual(m_, fld_->dev(), &r), "fdMatrix::residual");
        return r;
    }
    void setReference(int i, int j, int k, scalar val)
    {
        nova::check(nova_fd_set_reference(m_, i, j, k, val), "fdMatrix::setReference");
    }
    fdMatrix &operator+=(const fdMatrix &r)
    {
        nova::check(nova_fd_combine(m_, r.m_, 1.0), "fdMatrix::operator+=");
        return *this;
    }
    fdMatrix &operator-=(const fdMatrix &r)
    {
        nova::check(nova_fd_combine(m_, r.m_, -1.0), "fdMatrix::operator-=");
        return *this;
    }
    void negate() { nova::check(nova_fd_combine(m_, nullptr, -1.0), "fdMatrix::operator-"); }
    void addSource(const Field<T> &f) { nova::check(nova_fd_add_source(m_, f.dev()), "fdMatrix::operator=="); }
};

template <class T>
std::shared_ptr<fdMatrix<T>> laplacian(const scalar mu, std::shared_ptr<Field<T>> fld)
{
    nova_fdmatrix *m = nullptr;
    nova::check(nova_fd_laplacian(fld->mesh().ctx(), mu, fld->dev(), &m), "im::laplacian");
    return std::make_shared<fdMatrix<T>>(m, fld);
}
template <class T>
std::shared_ptr<fdMatrix<T>> laplacian(std::shared_ptr<Field<T>> fld) { return im::laplacian(scalar(1.0), fld); }
// global operators (fdMatrix.hxx:359-413): they act on the left operand
template <class T>
std::shared_ptr<fdMatrix<T>> operator==(std::shared_ptr<fdMatrix<T>> m, const std::shared_ptr<Field<T>> &f)
{
    m->addSource(*f);
    return m;
}
template <class T>
std::shared_ptr<fdMatrix<T>> operator+(std::shared_ptr<fdMatrix<T>> a, const std::shared_ptr<fdMatrix<T>> &b)
{
    *a += *b;
    return a;
}
template <class T>
std::shared_ptr<fdMatrix<T>> operator-(std::shared_ptr<fdMatrix<T>> a, const std::shared_ptr<fdMatrix<T>> &b)
{
    *a -= *b;
    return a;
}
template <class T>
std::shared_ptr<fdMatrix<T>> operator-(std::shared_ptr<fdMatrix<T>> a)
{
    a->negate();
    return a;
}
}  // namespace im

// ---- the HDF5 dumps of app/mhdjet (mhdjet.cpp:25-89, 277-400) ----
namespace nova {
// component comp of f as dataset /3DArray of file (N_x x N_y x N_z doubles)
template <class T>
void writeHDF5(const std::string &file, const Field<T> &f, int comp = 0)
{
    check(nova_hdf5_write(f.mesh().ctx(), file.c_str(), f.dev(), pTraits<T>::nComp, comp), "HDF5 write");
}
}  // namespace nova

// ---- tools (src/Tools/Tools.h:19-36, Tools.cpp:43-62) ----
namespace tools {
const scalar pi = 3.14159265358979323846;
const scalar eps = 1.0e-16;
// tools::CFL (src/Tools/Tools.cpp:9-41): a device reduction, printed as
// the reference prints it (its cfl_y and cfl_z take max(cfl_x, ...) at the
// last node, as written); the Courant numbers max |U_d| dt / h_d come back
// in the returned array's last three entries.
inline std::array<scalar, 6> CFL(const Field<vector> &U, const Mesh &mesh)
{
    std::array<scalar, 6> c{};
    nova::check(nova_cfl(mesh.ctx(), U.dev(), mesh.time().dt(), c.data()), "tools::CFL");
    std::cout << "================================" << std::endl;
    std::cout << "  CFL x: " << c[0] << ".\n";
    std::cout << "  CFL y: " << c[1] << ".\n";
    std::cout << "  CFL z: " << c[2] << ".\n";
    std::cout << "================================" << std::endl;
    return c;
}
inline scalar maximum(const FieldPtr<scalar> &f)
{
    // max |f| over [m/2, n - m/2) = every stored node
    std::vector<double> h((size_t)f->mesh().points());
    nova::check(nova_memcpy_d2h(f->mesh().ctx(), h.data(), f->dev(), h.size()), "maximum");
    scalar mx = 0.0;
    for (double v : h) mx = std::fabs(v) > mx ? std::fabs(v) : mx;
    return mx;
}
}  // namespace tools
