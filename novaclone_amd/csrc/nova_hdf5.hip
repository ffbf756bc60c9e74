// nova_hdf5.hip -- the HDF5 field dumps of the reference's mhdjet app.
//
// app/mhdjet/mhdjet.cpp:25-89 writes each field component to its own file:
// one dataset "/3DArray" of native doubles, dimensions (N_x, N_y, N_z) (the
// global cell counts), element [i][j][k] = the node (i, j, k) of the
// interior without the duplicated periodic endpoint (mhdjet.cpp:277-290,
// i slowest, k fastest).  That is the stored layout of a component here, so
// a dump is one copy to the host and one dataset write (the first N_d of
// the N_d + 1 stored nodes of a walled direction are selected by a
// hyperslab of the memory space).  One rank: no MPI-IO file access list.
//
// The HDF5 C library is opened at the first call (dlopen; the environment
// variable NOVA_HDF5_LIB, else libhdf5.so on the loader's path, else
// /opt/conda/lib/libhdf5.so), so the rest of the library loads and runs
// where HDF5 is absent; a call then fails with NOVA_ERR_UNSUPPORTED.  The
// 1.10 ABI is assumed (64-bit identifiers), and checked with
// H5get_libversion.
#include <dlfcn.h>

#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "nova_internal.hpp"

namespace nova {
namespace {

typedef int64_t hid_t;
typedef unsigned long long hsize_t;
typedef int herr_t;
constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;  // H5F_ACC_TRUNC, H5F_ACC_RDONLY
constexpr hid_t kDefault = 0;                                   // H5P_DEFAULT, H5S_ALL
constexpr int kSelectSet = 0;                                   // H5S_SELECT_SET

struct H5Api {
    void *lib = nullptr;
    herr_t (*open)(void);
    herr_t (*libversion)(unsigned *, unsigned *, unsigned *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
    int (*sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char *, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dclose)(hid_t);
    hid_t *native_double;  // H5T_NATIVE_DOUBLE_g, valid after H5open
};

std::mutex g_h5_mu;
H5Api g_h5;
std::string g_h5_err;

template <class F>
bool sym(void *lib, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

// Opens the library once; returns nullptr (and sets the reason) if it is
// absent or not a 1.10+ ABI.
H5Api *h5()
{
    std::lock_guard<std::mutex> lk(g_h5_mu);
    if (g_h5.lib) return &g_h5;
    if (!g_h5_err.empty()) return nullptr;
    std::vector<std::string> cand;
    if (const char *e = std::getenv("NOVA_HDF5_LIB")) cand.push_back(e);
    cand.push_back("libhdf5.so");
    cand.push_back("/opt/conda/lib/libhdf5.so");
    void *lib = nullptr;
    for (const auto &c : cand)
        if ((lib = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) {
        g_h5_err = "the HDF5 C library (libhdf5.so) was not found: set NOVA_HDF5_LIB";
        return nullptr;
    }
    H5Api a{};
    a.lib = lib;
    bool ok = sym(lib, "H5open", a.open) && sym(lib, "H5get_libversion", a.libversion) &&
              sym(lib, "H5Fcreate", a.fcreate) && sym(lib, "H5Fopen", a.fopen) && sym(lib, "H5Fclose", a.fclose) &&
              sym(lib, "H5Screate_simple", a.screate_simple) && sym(lib, "H5Sselect_hyperslab", a.sselect_hyperslab) &&
              sym(lib, "H5Sget_simple_extent_dims", a.sget_simple_extent_dims) && sym(lib, "H5Sclose", a.sclose) &&
              sym(lib, "H5Dcreate2", a.dcreate2) && sym(lib, "H5Dopen2", a.dopen2) &&
              sym(lib, "H5Dget_space", a.dget_space) && sym(lib, "H5Dwrite", a.dwrite) &&
              sym(lib, "H5Dread", a.dread) && sym(lib, "H5Dclose", a.dclose) &&
              sym(lib, "H5T_NATIVE_DOUBLE_g", a.native_double);
    unsigned maj = 0, min = 0, rel = 0;
    if (ok) ok = a.open() >= 0 && a.libversion(&maj, &min, &rel) >= 0 && (maj > 1 || (maj == 1 && min >= 10));
    if (!ok) {
        dlclose(lib);
        g_h5_err = "the HDF5 library found lacks the 1.10 C API";
        return nullptr;
    }
    g_h5 = a;
    return &g_h5;
}

}  // namespace

// dims: the dataset's extents; stored: the extents of the host array (each
// >= dims; the leading dims of each direction are written).
int hdf5_write_host(const char *file, const int64_t dims[3], const int64_t stored[3], const double *host)
{
    H5Api *H = h5();
    if (!H) return fail(NOVA_ERR_UNSUPPORTED, g_h5_err);
    for (int a = 0; a < 3; a++)
        if (dims[a] < 1 || stored[a] < dims[a]) return fail(NOVA_ERR_ARG, "HDF5 write: bad extents");
    std::lock_guard<std::mutex> lk(g_h5_mu);  // the serial library is not thread-safe
    const hsize_t d[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    const hsize_t s[3] = {(hsize_t)stored[0], (hsize_t)stored[1], (hsize_t)stored[2]};
    const hsize_t start[3] = {0, 0, 0};
    const hid_t f = H->fcreate(file, kAccTrunc, kDefault, kDefault);
    if (f < 0) return fail(NOVA_ERR_IO, std::string("HDF5: cannot create ") + file);
    const hid_t fs = H->screate_simple(3, d, nullptr);
    const hid_t ms = H->screate_simple(3, s, nullptr);
    const hid_t ds = H->dcreate2(f, "/3DArray", *H->native_double, fs, kDefault, kDefault, kDefault);
    herr_t e = (fs < 0 || ms < 0 || ds < 0) ? -1 : 0;
    if (e >= 0) e = H->sselect_hyperslab(ms, kSelectSet, start, nullptr, d, nullptr);
    if (e >= 0) e = H->dwrite(ds, *H->native_double, ms, fs, kDefault, host);
    if (ds >= 0) H->dclose(ds);
    if (ms >= 0) H->sclose(ms);
    if (fs >= 0) H->sclose(fs);
    const herr_t ec = H->fclose(f);
    if (e < 0 || ec < 0) return fail(NOVA_ERR_IO, std::string("HDF5: writing /3DArray of ") + file + " failed");
    return NOVA_OK;
}

int hdf5_read_host(const char *file, int64_t dims[3], double *host, size_t cap)
{
    H5Api *H = h5();
    if (!H) return fail(NOVA_ERR_UNSUPPORTED, g_h5_err);
    std::lock_guard<std::mutex> lk(g_h5_mu);
    const hid_t f = H->fopen(file, kAccRdonly, kDefault);
    if (f < 0) return fail(NOVA_ERR_IO, std::string("HDF5: cannot open ") + file);
    const hid_t ds = H->dopen2(f, "/3DArray", kDefault);
    int rc = NOVA_OK;
    hid_t sp = -1;
    hsize_t d[3] = {0, 0, 0};
    if (ds < 0) rc = fail(NOVA_ERR_IO, std::string("HDF5: no /3DArray in ") + file);
    if (!rc) {
        sp = H->dget_space(ds);
        if (sp < 0 || H->sget_simple_extent_dims(sp, d, nullptr) != 3)
            rc = fail(NOVA_ERR_IO, "HDF5: /3DArray is not three-dimensional");
    }
    if (!rc) {
        for (int a = 0; a < 3; a++) dims[a] = (int64_t)d[a];
        if (host) {
            if ((size_t)(d[0] * d[1] * d[2]) > cap) rc = fail(NOVA_ERR_ARG, "HDF5 read: buffer too small");
            else if (H->dread(ds, *H->native_double, kDefault, kDefault, kDefault, host) < 0)
                rc = fail(NOVA_ERR_IO, "HDF5: reading /3DArray failed");
        }
    }
    if (sp >= 0) H->sclose(sp);
    if (ds >= 0) H->dclose(ds);
    H->fclose(f);
    return rc;
}

// mhdjet.cpp:279-400: component comp of a stored field on the device.
int hdf5_write_field(nova_ctx *ctx, const char *file, const double *F, int comp)
{
    const Grid &g = ctx->g;
    std::vector<double> h((size_t)g.np);
    NOVA_HIP(hipMemcpyAsync(h.data(), F + (size_t)comp * g.np, sizeof(double) * g.np, hipMemcpyDeviceToHost,
                            ctx->stream));
    NOVA_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t dims[3] = {g.n[0], g.n[1], g.n[2]};
    const int64_t stored[3] = {g.s[0], g.s[1], g.s[2]};
    return hdf5_write_host(file, dims, stored, h.data());
}

}  // namespace nova
