// kwh_h5.cpp -- k-Wave HDF5 input / output for libkwavehip (host code only).
//
// The reference writes a k-Wave C++ input file and reads the engine's p_final through its CHDF5
// binding (Sources/KWave/IO/KWaveInput.swift:12-99, Sources/KWave/IO/HDF5.swift:26-190). This file
// reads and writes the same format so that a k-Wave input file can be run on the device
// (kwh_h5_read_input -> kwh_sim_create -> kwh_sim_run -> kwh_h5_write_output) and an API-built
// simulation can be saved as one (kwh_h5_write_input).
//
// libhdf5 is not linked: it is loaded with dlopen on first use (KWH_HDF5_LIBRARY, then the usual
// sonames, then /opt/conda/lib), so the library itself carries no HDF5 dependency. Every entry
// point returns KWH_EUNSUPPORTED with a message when no libhdf5 >= 1.10 can be loaded.
//
// Format (as the reference writes it, KWaveInput.swift:33-98, HDF5.swift:103-147):
//  - root string attributes created_by, creation_date, file_description, file_type
//    ("input" / "output"), major_version "1", minor_version "2";
//  - integer scalars (Nx, Ny, Nz, Nt, the *_flag values, sensor_mask_type, pml_*_size) as uint64
//    datasets of shape [1, 1, 1]; real scalars (dx, dy, dt, c0, c_ref, rho0, rho0_sg*, pml_*_alpha)
//    as float32 [1, 1, 1]; z-variables only in 3D;
//  - fields as float32 [Nz, Ny, Nx] with x fastest (the transpose of the C-order [nx][ny][nz]
//    arrays of the API);
//  - sensor_mask_index as uint64 [1, 1, Ns]: 1-based column-major (x fastest) linear indices of
//    the mask points, ascending;
//  - every dataset carries string attributes domain_type "real" and data_type "float" / "long".
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kwave_hip.h"

// Defined with the rest of the C ABI (kwh_sim.hip), not exported: the thread's last-error message
// (kwh_last_error) and the element size of a sim's fields (4 or 8).
void kwh_set_last_error(const char* msg);
int kwh_sim_elem_bytes(const kwh_sim* sim);

namespace {

struct H5Error : std::runtime_error {
    int code;
    H5Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define H5_REQUIRE(cond, code, msg)                           \
    do {                                                      \
        if (!(cond)) throw H5Error((code), std::string(msg)); \
    } while (0)

template <class F>
int guarded_h5(F&& f) {
    try {
        f();
        kwh_set_last_error("");
        return KWH_OK;
    } catch (const H5Error& e) {
        kwh_set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        kwh_set_last_error("host allocation failed");
        return KWH_ENOMEM;
    } catch (const std::exception& e) {
        kwh_set_last_error(e.what());
        return KWH_EINVAL;
    }
}

// ------------------------------------------------------------------ libhdf5, loaded at run time
// The C types and constants of the HDF5 >= 1.10 ABI that this file uses.
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;          // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr unsigned kAccRdonly = 0u;    // H5F_ACC_RDONLY
constexpr unsigned kAccTrunc = 2u;     // H5F_ACC_TRUNC
constexpr int kScalarSpace = 0;        // H5S_SCALAR
constexpr int kStrNullTerm = 0;        // H5T_STR_NULLTERM
constexpr int kClassString = 3;        // H5T_STRING

struct Hdf5 {
    void* so = nullptr;
    std::string error;
    herr_t (*open)();
    herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*);
    herr_t (*Eset_auto2)(hid_t, void*, void*);
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    herr_t (*Dclose)(hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    hid_t (*Dget_space)(hid_t);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    hid_t (*Screate)(int);
    herr_t (*Sclose)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tset_strpad)(hid_t, int);
    size_t (*Tget_size)(hid_t);
    int (*Tget_class)(hid_t);
    herr_t (*Tclose)(hid_t);
    hid_t (*Acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Aopen)(hid_t, const char*, hid_t);
    herr_t (*Awrite)(hid_t, hid_t, const void*);
    herr_t (*Aread)(hid_t, hid_t, void*);
    hid_t (*Aget_type)(hid_t);
    herr_t (*Aclose)(hid_t);
    htri_t (*Aexists)(hid_t, const char*);
    htri_t (*Lexists)(hid_t, const char*, hid_t);
    // native type ids (the *_g globals, valid after H5open)
    hid_t t_float = -1, t_double = -1, t_uint64 = -1, t_c_s1 = -1;
};

std::mutex g_h5_mu;   // libhdf5 builds are not thread-safe by default: one call at a time

template <class P>
bool bind(void* so, const char* name, P& fn) {
    fn = reinterpret_cast<P>(dlsym(so, name));
    return fn != nullptr;
}

Hdf5& hdf5() {
    static Hdf5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> names;
        if (const char* env = std::getenv("KWH_HDF5_LIBRARY")) names.push_back(env);
        for (const char* n : {"libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so.103", "libhdf5.so.200",
                              "libhdf5.so.310", "libhdf5_serial.so.200", "/opt/conda/lib/libhdf5.so.103",
                              "/opt/conda/lib/libhdf5.so"})
            names.push_back(n);
        for (const auto& n : names) {
            h.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (h.so) break;
        }
        if (!h.so) {
            h.error = "libhdf5 not found (set KWH_HDF5_LIBRARY to its path)";
            return;
        }
        bool ok = bind(h.so, "H5open", h.open) && bind(h.so, "H5get_libversion", h.get_libversion) &&
                  bind(h.so, "H5Eset_auto2", h.Eset_auto2) && bind(h.so, "H5Fcreate", h.Fcreate) &&
                  bind(h.so, "H5Fopen", h.Fopen) && bind(h.so, "H5Fclose", h.Fclose) &&
                  bind(h.so, "H5Dcreate2", h.Dcreate2) && bind(h.so, "H5Dopen2", h.Dopen2) &&
                  bind(h.so, "H5Dclose", h.Dclose) && bind(h.so, "H5Dread", h.Dread) &&
                  bind(h.so, "H5Dwrite", h.Dwrite) && bind(h.so, "H5Dget_space", h.Dget_space) &&
                  bind(h.so, "H5Sget_simple_extent_ndims", h.Sget_simple_extent_ndims) &&
                  bind(h.so, "H5Sget_simple_extent_dims", h.Sget_simple_extent_dims) &&
                  bind(h.so, "H5Screate_simple", h.Screate_simple) && bind(h.so, "H5Screate", h.Screate) &&
                  bind(h.so, "H5Sclose", h.Sclose) && bind(h.so, "H5Tcopy", h.Tcopy) &&
                  bind(h.so, "H5Tset_size", h.Tset_size) && bind(h.so, "H5Tset_strpad", h.Tset_strpad) &&
                  bind(h.so, "H5Tget_size", h.Tget_size) && bind(h.so, "H5Tget_class", h.Tget_class) &&
                  bind(h.so, "H5Tclose", h.Tclose) && bind(h.so, "H5Acreate2", h.Acreate2) &&
                  bind(h.so, "H5Aopen", h.Aopen) && bind(h.so, "H5Awrite", h.Awrite) &&
                  bind(h.so, "H5Aread", h.Aread) && bind(h.so, "H5Aget_type", h.Aget_type) &&
                  bind(h.so, "H5Aclose", h.Aclose) && bind(h.so, "H5Aexists", h.Aexists) &&
                  bind(h.so, "H5Lexists", h.Lexists);
        if (!ok) {
            h.error = "libhdf5 lacks a required symbol";
            return;
        }
        unsigned maj = 0, min = 0, rel = 0;
        if (h.open() < 0 || h.get_libversion(&maj, &min, &rel) < 0 || maj < 1 || (maj == 1 && min < 10)) {
            h.error = "libhdf5 >= 1.10 required (64-bit identifiers)";
            return;
        }
        auto global = [&](const char* name, hid_t& out) {
            const hid_t* p = reinterpret_cast<const hid_t*>(dlsym(h.so, name));
            if (p) out = *p;
            return p != nullptr && out >= 0;
        };
        if (!(global("H5T_NATIVE_FLOAT_g", h.t_float) && global("H5T_NATIVE_DOUBLE_g", h.t_double) &&
              global("H5T_NATIVE_UINT64_g", h.t_uint64) && global("H5T_C_S1_g", h.t_c_s1))) {
            h.error = "libhdf5 native type identifiers unavailable";
            return;
        }
        h.Eset_auto2(kDefault, nullptr, nullptr);   // errors come back as return codes, not prints
    });
    H5_REQUIRE(h.error.empty(), KWH_EUNSUPPORTED, "k-Wave HDF5 I/O unavailable: " + h.error);
    return h;
}

// RAII identifier
struct Id {
    hid_t v = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t x, herr_t (*c)(hid_t)) : v(x), close(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id() {
        if (v >= 0 && close) close(v);
    }
    operator hid_t() const { return v; }
};

// ------------------------------------------------------------------ writing
struct Writer {
    Hdf5& h;
    hid_t file;

    void string_attr(hid_t obj, const char* name, const std::string& value) {
        Id type(h.Tcopy(h.t_c_s1), h.Tclose);
        H5_REQUIRE(type >= 0 && h.Tset_size(type, value.size() + 1) >= 0 && h.Tset_strpad(type, kStrNullTerm) >= 0,
                   KWH_EINVAL, std::string("HDF5: string type for attribute ") + name);
        Id space(h.Screate(kScalarSpace), h.Sclose);
        Id attr(h.Acreate2(obj, name, type, space, kDefault, kDefault), h.Aclose);
        H5_REQUIRE(attr >= 0 && h.Awrite(attr, type, value.c_str()) >= 0, KWH_EINVAL,
                   std::string("HDF5: failed to write attribute ") + name);
    }
    void root_attrs(const char* file_type, const char* description) {
        char date[64];
        std::time_t now = std::time(nullptr);
        std::tm tmv{};
        gmtime_r(&now, &tmv);
        std::strftime(date, sizeof(date), "%d-%b-%Y-%H-%M-%S", &tmv);
        string_attr(file, "created_by", "kwavehip");
        string_attr(file, "creation_date", date);
        string_attr(file, "file_description", description);
        string_attr(file, "file_type", file_type);
        string_attr(file, "major_version", "1");
        string_attr(file, "minor_version", "2");
    }
    void dataset(const char* name, const std::vector<hsize_t>& dims, hid_t type, const void* data,
                 const char* data_type) {
        Id space(h.Screate_simple((int)dims.size(), dims.data(), nullptr), h.Sclose);
        Id d(h.Dcreate2(file, name, type, space, kDefault, kDefault, kDefault), h.Dclose);
        H5_REQUIRE(space >= 0 && d >= 0 && h.Dwrite(d, type, kDefault, kDefault, kDefault, data) >= 0, KWH_EINVAL,
                   std::string("HDF5: failed to write dataset ") + name);
        string_attr(d, "domain_type", "real");
        string_attr(d, "data_type", data_type);
    }
    void u64(const char* name, uint64_t v) { dataset(name, {1, 1, 1}, h.t_uint64, &v, "long"); }
    void f32(const char* name, double v) {
        const float f = (float)v;
        dataset(name, {1, 1, 1}, h.t_float, &f, "float");
    }
};

// C-order [nx][ny][nz] -> disk order [Nz][Ny][Nx] (x fastest), and back.
template <class T, class U>
void to_disk(const T* c, U* disk, int64_t nx, int64_t ny, int64_t nz) {
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) disk[(z * ny + y) * nx + x] = (U)c[(x * ny + y) * nz + z];
}
template <class T, class U>
void from_disk(const T* disk, U* c, int64_t nx, int64_t ny, int64_t nz) {
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t z = 0; z < nz; ++z) c[(x * ny + y) * nz + z] = (U)disk[(z * ny + y) * nx + x];
}

// rho0 at +d/2 along `axis` (neighbour mean, the last cell replicated; staggerDensity,
// KSpaceFirstOrder.swift:799-806), C-order.
std::vector<float> staggered(const float* r, const int64_t n[3], int axis) {
    const int64_t N = n[0] * n[1] * n[2];
    std::vector<float> out(N);
    const int64_t stride = axis == 0 ? n[1] * n[2] : (axis == 1 ? n[2] : 1);
    for (int64_t i = 0; i < N; ++i) {
        const int64_t c = (i / stride) % n[axis];
        out[i] = c + 1 < n[axis] ? 0.5f * (r[i] + r[i + stride]) : r[i];
    }
    return out;
}

// ------------------------------------------------------------------ reading
struct Reader {
    Hdf5& h;
    hid_t file;

    bool exists(const char* name) { return h.Lexists(file, name, kDefault) > 0; }
    std::vector<hsize_t> dims(const char* name) {
        Id d(h.Dopen2(file, name, kDefault), h.Dclose);
        H5_REQUIRE(d >= 0, KWH_EINVAL, std::string("HDF5: dataset ") + name + " not found");
        Id s(h.Dget_space(d), h.Sclose);
        const int nd = h.Sget_simple_extent_ndims(s);
        H5_REQUIRE(nd >= 0 && nd <= 8, KWH_EINVAL, std::string("HDF5: bad dataspace of ") + name);
        std::vector<hsize_t> dm(std::max(nd, 0));
        if (nd > 0) h.Sget_simple_extent_dims(s, dm.data(), nullptr);
        return dm;
    }
    template <class T>
    std::vector<T> read(const char* name, hid_t mem_type) {
        const auto dm = dims(name);
        size_t count = 1;
        for (auto v : dm) count *= (size_t)v;
        std::vector<T> out(count);
        Id d(h.Dopen2(file, name, kDefault), h.Dclose);
        H5_REQUIRE(h.Dread(d, mem_type, kDefault, kDefault, kDefault, out.data()) >= 0, KWH_EINVAL,
                   std::string("HDF5: failed to read dataset ") + name);
        return out;
    }
    double scalar(const char* name) {
        auto v = read<double>(name, h.t_double);
        H5_REQUIRE(v.size() == 1, KWH_EINVAL, std::string("HDF5: ") + name + " is not a scalar");
        return v[0];
    }
    double scalar_or(const char* name, double dflt) { return exists(name) ? scalar(name) : dflt; }
    uint64_t flag(const char* name) { return exists(name) ? (uint64_t)std::llround(scalar(name)) : 0; }
    std::string root_attr(const char* name) {
        if (h.Aexists(file, name) <= 0) return std::string();
        Id a(h.Aopen(file, name, kDefault), h.Aclose);
        Id t(h.Aget_type(a), h.Tclose);
        H5_REQUIRE(a >= 0 && t >= 0 && h.Tget_class(t) == kClassString, KWH_EINVAL,
                   std::string("HDF5: root attribute ") + name + " is not a string");
        std::vector<char> buf(h.Tget_size(t) + 1, 0);
        H5_REQUIRE(h.Aread(a, t, buf.data()) >= 0, KWH_EINVAL, std::string("HDF5: failed to read attribute ") + name);
        return std::string(buf.data());
    }
};

}  // namespace

// A k-Wave input file read into host memory: the descriptor points into these buffers.
struct kwh_h5_input {
    kwh_sim_desc desc{};
    std::vector<float> c0, rho0, bona, p0;
    std::vector<int64_t> sensor;       // ascending C-order flat indices
    std::vector<int64_t> file_order;   // file_order[k]: row of the k-th file index in `sensor`
};

extern "C" {

int kwh_h5_available(void) {
    return guarded_h5([] {
        std::lock_guard<std::mutex> lock(g_h5_mu);
        hdf5();
    });
}

int kwh_h5_write_input(const char* path, const kwh_sim_desc* d) {
    return guarded_h5([&] {
        H5_REQUIRE(path && d, KWH_EINVAL, "NULL argument");
        H5_REQUIRE(d->abi_version == KWH_ABI_VERSION, KWH_EINVAL, "descriptor ABI version mismatch");
        H5_REQUIRE(d->dim == 2 || d->dim == 3, KWH_EUNSUPPORTED, "the k-Wave input format holds 2D and 3D grids");
        const int64_t n[3] = {d->n[0], d->n[1], d->dim == 3 ? d->n[2] : 1};
        H5_REQUIRE(n[0] > 0 && n[1] > 0 && n[2] > 0 && d->nt > 0 && d->dt > 0, KWH_EINVAL, "grid sizes, nt and dt must be positive");
        const int64_t N = n[0] * n[1] * n[2];
        H5_REQUIRE(d->p0.data && d->p0.count == N, KWH_EUNSUPPORTED, "the k-Wave input writer needs a grid-shaped p0 source");
        H5_REQUIRE(d->sensor_cartesian == nullptr && (d->sensor_index != nullptr || d->n_sensor == 0), KWH_EUNSUPPORTED,
                   "the k-Wave input writer needs a binary sensor mask");
        H5_REQUIRE(d->directivity_angle == nullptr, KWH_EUNSUPPORTED, "directivity is not part of the k-Wave input format");
        H5_REQUIRE(!d->p_source.index && !d->u_source[0].index && !d->u_source[1].index && !d->u_source[2].index,
                   KWH_EUNSUPPORTED, "time-varying sources are not written (p0 problems only)");
        H5_REQUIRE(!d->alpha_coeff.data || d->alpha_mode == KWH_ALPHA_NONE, KWH_EUNSUPPORTED,
                   "absorbing media are stored as precomputed operators in the k-Wave format and are not written");
        auto field_ok = [&](const kwh_field& f) { return f.data && (f.count == 1 || f.count == N); };
        H5_REQUIRE(field_ok(d->sound_speed) && field_ok(d->density), KWH_EINVAL,
                   "sound speed and density must be scalars or grid-shaped");
        bool nonlinear = false;
        if (d->bona.data) {
            H5_REQUIRE(field_ok(d->bona), KWH_EINVAL, "B/A must be a scalar or grid-shaped");
            for (int64_t i = 0; i < d->bona.count; ++i) nonlinear |= d->bona.data[i] != 0.0f;
        }
        for (int64_t k = 0; k < d->n_sensor; ++k)
            H5_REQUIRE(d->sensor_index[k] >= 0 && d->sensor_index[k] < N && (k == 0 || d->sensor_index[k] > d->sensor_index[k - 1]),
                       KWH_EINVAL, "sensor indices must be ascending and inside the grid");

        std::lock_guard<std::mutex> lock(g_h5_mu);
        Hdf5& h = hdf5();
        Id file(h.Fcreate(path, kAccTrunc, kDefault, kDefault), h.Fclose);
        H5_REQUIRE(file >= 0, KWH_EINVAL, std::string("HDF5: could not create ") + path);
        Writer w{h, file};
        w.root_attrs("input", "k-Wave input file (kwavehip)");
        const bool is3 = d->dim == 3;
        w.u64("Nx", n[0]);
        w.u64("Ny", n[1]);
        w.u64("Nz", n[2]);
        w.u64("Nt", (uint64_t)d->nt);
        w.u64("p0_source_flag", 1);
        for (const char* f : {"p_source_flag", "ux_source_flag", "uy_source_flag", "uz_source_flag", "sxx_source_flag",
                              "syy_source_flag", "szz_source_flag", "sxy_source_flag", "sxz_source_flag",
                              "syz_source_flag", "transducer_source_flag", "nonuniform_grid_flag"})
            w.u64(f, 0);
        w.u64("nonlinear_flag", nonlinear ? 1 : 0);
        w.u64("absorbing_flag", 0);
        w.u64("elastic_flag", 0);
        w.u64("axisymmetric_flag", 0);
        w.u64("sensor_mask_type", 0);
        w.u64("pml_x_size", (uint64_t)d->pml_size);
        w.u64("pml_y_size", (uint64_t)d->pml_size);
        w.u64("pml_z_size", is3 ? (uint64_t)d->pml_size : 0);
        w.f32("dx", d->d[0]);
        w.f32("dy", d->d[1]);
        if (is3) w.f32("dz", d->d[2]);
        w.f32("pml_x_alpha", d->pml_alpha);
        w.f32("pml_y_alpha", d->pml_alpha);
        if (is3) w.f32("pml_z_alpha", d->pml_alpha);
        w.f32("dt", d->dt);
        const std::vector<hsize_t> fdims = {(hsize_t)n[2], (hsize_t)n[1], (hsize_t)n[0]};
        std::vector<float> disk(N);
        auto write_field = [&](const char* name, const float* c) {
            to_disk(c, disk.data(), n[0], n[1], n[2]);
            w.dataset(name, fdims, h.t_float, disk.data(), "float");
        };
        auto write_map = [&](const char* name, const kwh_field& f) {
            if (f.count == 1) w.f32(name, f.data[0]);
            else write_field(name, f.data);
        };
        write_map("c0", d->sound_speed);
        float cmax = d->sound_speed.data[0];
        for (int64_t i = 1; i < d->sound_speed.count; ++i) cmax = std::max(cmax, d->sound_speed.data[i]);
        w.f32("c_ref", cmax);                      // c_ref = max(c0), as the solver takes it
        write_map("rho0", d->density);
        const char* sg[3] = {"rho0_sgx", "rho0_sgy", "rho0_sgz"};
        for (int a = 0; a < d->dim; ++a) {
            if (d->density.count == 1) w.f32(sg[a], d->density.data[0]);
            else write_field(sg[a], staggered(d->density.data, n, a).data());
        }
        if (nonlinear) write_map("BonA", d->bona);
        write_field("p0_source_input", d->p0.data);
        // 1-based column-major indices, ascending
        std::vector<uint64_t> idx((size_t)d->n_sensor);
        for (int64_t k = 0; k < d->n_sensor; ++k) {
            const int64_t c = d->sensor_index[k];
            const int64_t x = c / (n[1] * n[2]), y = (c / n[2]) % n[1], z = c % n[2];
            idx[k] = (uint64_t)(x + n[0] * (y + n[1] * z) + 1);
        }
        std::sort(idx.begin(), idx.end());
        w.dataset("sensor_mask_index", {1, 1, (hsize_t)idx.size()}, h.t_uint64, idx.data(), "long");
    });
}

int kwh_h5_read_input(const char* path, kwh_h5_input** out) {
    return guarded_h5([&] {
        H5_REQUIRE(path && out, KWH_EINVAL, "NULL argument");
        *out = nullptr;
        std::unique_ptr<kwh_h5_input> in(new kwh_h5_input());
        std::lock_guard<std::mutex> lock(g_h5_mu);
        Hdf5& h = hdf5();
        Id file(h.Fopen(path, kAccRdonly, kDefault), h.Fclose);
        H5_REQUIRE(file >= 0, KWH_EINVAL, std::string("HDF5: could not open ") + path);
        Reader r{h, file};
        const std::string type = r.root_attr("file_type");
        H5_REQUIRE(type.empty() || type == "input", KWH_EINVAL, "not a k-Wave input file (file_type " + type + ")");
        for (const char* f : {"p_source_flag", "ux_source_flag", "uy_source_flag", "uz_source_flag", "sxx_source_flag",
                              "syy_source_flag", "szz_source_flag", "sxy_source_flag", "sxz_source_flag",
                              "syz_source_flag", "transducer_source_flag", "nonuniform_grid_flag", "absorbing_flag",
                              "elastic_flag", "axisymmetric_flag", "sensor_mask_type"})
            H5_REQUIRE(r.flag(f) == 0, KWH_EUNSUPPORTED, std::string("k-Wave input with ") + f +
                                                              " set: only p0 problems with an index sensor mask "
                                                              "on lossless media are read");
        H5_REQUIRE(r.flag("p0_source_flag") == 1 && r.exists("p0_source_input"), KWH_EUNSUPPORTED,
                   "k-Wave input without a p0 source");
        kwh_sim_desc& d = in->desc;
        d.abi_version = KWH_ABI_VERSION;
        const double nxd = r.scalar("Nx"), nyd = r.scalar("Ny"), nzd = r.scalar_or("Nz", 1.0);
        H5_REQUIRE(nxd >= 1 && nyd >= 1 && nzd >= 1 && nxd * nyd * nzd < 9.2e18, KWH_EINVAL, "bad grid size");
        const int64_t n[3] = {(int64_t)nxd, (int64_t)nyd, (int64_t)nzd};
        H5_REQUIRE(n[1] > 1, KWH_EUNSUPPORTED, "1D k-Wave input files are not read");
        d.dim = n[2] > 1 ? 3 : 2;
        for (int a = 0; a < 3; ++a) d.n[a] = (int32_t)n[a];
        const int64_t N = n[0] * n[1] * n[2];
        d.d[0] = r.scalar("dx");
        d.d[1] = r.scalar("dy");
        d.d[2] = d.dim == 3 ? r.scalar("dz") : 1.0;
        d.dt = r.scalar("dt");
        d.nt = (int32_t)r.scalar("Nt");
        d.dtype = KWH_FLOAT32;
        const uint64_t px = r.flag("pml_x_size"), py = r.flag("pml_y_size"), pz = r.flag("pml_z_size");
        H5_REQUIRE(px == py && (d.dim == 2 || pz == px), KWH_EUNSUPPORTED, "the solver takes one PML size for every axis");
        d.pml_size = (int32_t)px;
        const double ax = r.scalar_or("pml_x_alpha", 2.0), ay = r.scalar_or("pml_y_alpha", ax);
        const double az = d.dim == 3 ? r.scalar_or("pml_z_alpha", ax) : ax;
        H5_REQUIRE(ax == ay && ax == az, KWH_EUNSUPPORTED, "the solver takes one PML alpha for every axis");
        d.pml_alpha = ax;
        d.smooth_p0 = d.smooth_c0 = d.smooth_rho0 = 0;   // the file holds the final fields
        auto read_map = [&](const char* name, std::vector<float>& dst, kwh_field& f) {
            const auto raw = r.read<float>(name, h.t_float);
            H5_REQUIRE(raw.size() == 1 || (int64_t)raw.size() == N, KWH_EINVAL,
                       std::string("HDF5: ") + name + " is neither a scalar nor grid-shaped");
            dst.resize(raw.size());
            if (raw.size() == 1) dst[0] = raw[0];
            else from_disk(raw.data(), dst.data(), n[0], n[1], n[2]);
            f.data = dst.data();
            f.count = (int64_t)dst.size();
        };
        read_map("c0", in->c0, d.sound_speed);
        read_map("rho0", in->rho0, d.density);
        if (r.flag("nonlinear_flag")) read_map("BonA", in->bona, d.bona);
        read_map("p0_source_input", in->p0, d.p0);
        H5_REQUIRE(d.p0.count == N, KWH_EINVAL, "p0_source_input is not grid-shaped");
        d.alpha_mode = KWH_ALPHA_NONE;
        d.alpha_power = 1.5;
        // sensor: 1-based column-major -> ascending C-order, remembering the file's order
        const auto idx = r.exists("sensor_mask_index") ? r.read<uint64_t>("sensor_mask_index", h.t_uint64)
                                                       : std::vector<uint64_t>();
        std::vector<std::pair<int64_t, int64_t>> cs(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            H5_REQUIRE(idx[k] >= 1 && (int64_t)idx[k] <= N, KWH_EINVAL, "sensor_mask_index outside the grid");
            const int64_t m = (int64_t)idx[k] - 1;
            const int64_t x = m % n[0], y = (m / n[0]) % n[1], z = m / (n[0] * n[1]);
            cs[k] = {(x * n[1] + y) * n[2] + z, (int64_t)k};
        }
        std::sort(cs.begin(), cs.end());
        in->sensor.resize(cs.size());
        in->file_order.resize(cs.size());
        for (size_t j = 0; j < cs.size(); ++j) {
            H5_REQUIRE(j == 0 || cs[j].first != cs[j - 1].first, KWH_EINVAL, "sensor_mask_index holds a point twice");
            in->sensor[j] = cs[j].first;
            in->file_order[cs[j].second] = (int64_t)j;
        }
        d.sensor_index = in->sensor.empty() ? nullptr : in->sensor.data();
        d.n_sensor = (int64_t)in->sensor.size();
        d.record = KWH_REC_P_FINAL;
        *out = in.release();
    });
}

const kwh_sim_desc* kwh_h5_input_desc(kwh_h5_input* in) { return in ? &in->desc : nullptr; }

void kwh_h5_input_destroy(kwh_h5_input* in) { delete in; }

int kwh_h5_write_output(const char* path, kwh_sim* sim, const kwh_h5_input* in) {
    return guarded_h5([&] {
        H5_REQUIRE(path && sim, KWH_EINVAL, "NULL argument");
        int64_t shape[3];
        int32_t nd = 0;
        H5_REQUIRE(kwh_sim_output_shape(sim, KWH_OUT_P_FINAL, shape, &nd) == KWH_OK, KWH_EINVAL, kwh_last_error());
        H5_REQUIRE(nd == 2 || nd == 3, KWH_EUNSUPPORTED, "k-Wave output files hold 2D and 3D grids");
        const int64_t n[3] = {shape[0], shape[1], nd == 3 ? shape[2] : 1};
        const int64_t N = n[0] * n[1] * n[2];
        const bool f32 = kwh_sim_elem_bytes(sim) == 4;
        std::vector<float> buf32(f32 ? N : 0);
        std::vector<double> buf64(f32 ? 0 : N);
        int rc = f32 ? kwh_sim_read_output(sim, KWH_OUT_P_FINAL, buf32.data(), N * sizeof(float))
                     : kwh_sim_read_output(sim, KWH_OUT_P_FINAL, buf64.data(), N * sizeof(double));
        H5_REQUIRE(rc == KWH_OK, rc, kwh_last_error());
        std::lock_guard<std::mutex> lock(g_h5_mu);
        Hdf5& h = hdf5();
        Id file(h.Fcreate(path, kAccTrunc, kDefault, kDefault), h.Fclose);
        H5_REQUIRE(file >= 0, KWH_EINVAL, std::string("HDF5: could not create ") + path);
        Writer w{h, file};
        w.root_attrs("output", "k-Wave output file (kwavehip)");
        const hid_t t = f32 ? h.t_float : h.t_double;
        const char* tn = f32 ? "float" : "double";
        const std::vector<hsize_t> fdims = {(hsize_t)n[2], (hsize_t)n[1], (hsize_t)n[0]};
        if (f32) {
            std::vector<float> disk(N);
            to_disk(buf32.data(), disk.data(), n[0], n[1], n[2]);
            w.dataset("p_final", fdims, t, disk.data(), tn);
        } else {
            std::vector<double> disk(N);
            to_disk(buf64.data(), disk.data(), n[0], n[1], n[2]);
            w.dataset("p_final", fdims, t, disk.data(), tn);
        }
        // recorded sensor fields, rows in the input file's sensor order: p as [1, Nt, Ns] (sensor
        // fastest), the aggregates as [1, 1, Ns]
        struct F { int out; const char* name; };
        for (F f : {F{KWH_OUT_P, "p"}, F{KWH_OUT_P_MAX, "p_max"}, F{KWH_OUT_P_MIN, "p_min"}, F{KWH_OUT_P_RMS, "p_rms"}}) {
            int64_t sh[3];
            int32_t snd = 0;
            if (kwh_sim_output_shape(sim, f.out, sh, &snd) != KWH_OK) continue;
            const int64_t ns = sh[0], steps = snd == 2 ? sh[1] : 1;
            if (ns == 0) continue;
            const size_t es = f32 ? 4 : 8;
            std::vector<char> raw((size_t)(ns * steps) * es), disk(raw.size());
            rc = kwh_sim_read_output(sim, f.out, raw.data(), raw.size());
            H5_REQUIRE(rc == KWH_OK, rc, kwh_last_error());
            H5_REQUIRE(!in || (int64_t)in->file_order.size() == ns, KWH_EINVAL, "input file and sim differ in sensor points");
            for (int64_t k = 0; k < ns; ++k) {
                const int64_t row = in ? in->file_order[k] : k;   // file position k holds sim row `row`
                for (int64_t t_ = 0; t_ < steps; ++t_)
                    std::memcpy(&disk[(size_t)(t_ * ns + k) * es], &raw[(size_t)(row * steps + t_) * es], es);
            }
            w.dataset(f.name, {1, (hsize_t)steps, (hsize_t)ns}, t, disk.data(), tn);
        }
    });
}

int kwh_h5_read_dataset(const char* path, const char* name, int64_t dims[8], int32_t* ndim, double* dst,
                        size_t dst_count) {
    return guarded_h5([&] {
        H5_REQUIRE(path && name && dims && ndim, KWH_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lock(g_h5_mu);
        Hdf5& h = hdf5();
        Id file(h.Fopen(path, kAccRdonly, kDefault), h.Fclose);
        H5_REQUIRE(file >= 0, KWH_EINVAL, std::string("HDF5: could not open ") + path);
        Reader r{h, file};
        H5_REQUIRE(r.exists(name), KWH_EINVAL, std::string("HDF5: dataset ") + name + " not found");
        const auto dm = r.dims(name);
        *ndim = (int32_t)dm.size();
        size_t count = 1;
        for (size_t i = 0; i < dm.size(); ++i) {
            dims[i] = (int64_t)dm[i];
            count *= (size_t)dm[i];
        }
        if (!dst) return;   // shape query
        H5_REQUIRE(dst_count >= count, KWH_EINVAL, "destination buffer too small");
        const auto v = r.read<double>(name, h.t_double);
        std::copy(v.begin(), v.end(), dst);
    });
}

int kwh_h5_read_root_attribute(const char* path, const char* name, char* dst, size_t dst_bytes) {
    return guarded_h5([&] {
        H5_REQUIRE(path && name && dst && dst_bytes > 0, KWH_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lock(g_h5_mu);
        Hdf5& h = hdf5();
        Id file(h.Fopen(path, kAccRdonly, kDefault), h.Fclose);
        H5_REQUIRE(file >= 0, KWH_EINVAL, std::string("HDF5: could not open ") + path);
        Reader r{h, file};
        H5_REQUIRE(h.Aexists(file, name) > 0, KWH_EINVAL, std::string("HDF5: root attribute ") + name + " not found");
        const std::string v = r.root_attr(name);
        H5_REQUIRE(v.size() < dst_bytes, KWH_EINVAL, "destination buffer too small");
        std::memcpy(dst, v.c_str(), v.size() + 1);
    });
}

}  // extern "C"
