// sph_openpmd.cpp -- openPMD 1.1.0 particle output over the HDF5 C API (libspearhed_openpmd.so).
//
// Replaces the reference's openPMD plugin (include/spearhed/plugins/openPMD/OpenPMDOutput.hpp:45-166,
// OpenPMDPlugin.x.cpp notify), which writes through openPMD-api: one file per iteration, records
// named and dimensioned as in include/spearhed/plugins/openPMD/ParticleTraits.hpp. The layout below
// is the openPMD standard's (base standard 1.1.0):
//   /                                  openPMD, openPMDextension, basePath, particlesPath,
//                                      iterationEncoding, iterationFormat, author, software,
//                                      softwareVersion, date
//   /data/<it>/                        time, dt, timeUnitSI
//   /data/<it>/particles/<species>/
//       <record>/                      unitDimension (7 x double), timeOffset (float)
//           x, y, z                    datasets [N] with unitSI (double)
//       <scalar record>                dataset [N] with unitDimension, timeOffset, unitSI
//       positionOffset/x, y, z         constant components: groups with value (double) and shape
// Host code only; the step library is used for sph_download / sph_reduce_conserved.
#include <hdf5.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/spearhed_openpmd.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// RAII holder of one HDF5 identifier
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

struct Writer {
    bool bad = false;
    std::string why;

    void check(bool good, const char* what)
    {
        if (!good && !bad) {
            bad = true;
            why = what;
        }
    }

    void attr_string(hid_t obj, const char* name, const std::string& v)
    {
        Hid type(H5Tcopy(H5T_C_S1), H5Tclose);
        check(type.ok(), "H5Tcopy");
        if (bad) return;
        check(H5Tset_size(type.id, v.size() + 1) >= 0, "H5Tset_size");
        check(H5Tset_strpad(type.id, H5T_STR_NULLTERM) >= 0, "H5Tset_strpad");
        Hid space(H5Screate(H5S_SCALAR), H5Sclose);
        Hid a(H5Acreate2(obj, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        check(a.ok(), name);
        if (!bad) check(H5Awrite(a.id, type.id, v.c_str()) >= 0, name);
    }

    template <class T>
    void attr_scalar(hid_t obj, const char* name, hid_t mem_type, T v)
    {
        Hid space(H5Screate(H5S_SCALAR), H5Sclose);
        Hid a(H5Acreate2(obj, name, mem_type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        check(a.ok(), name);
        if (!bad) check(H5Awrite(a.id, mem_type, &v) >= 0, name);
    }

    template <class T>
    void attr_array(hid_t obj, const char* name, hid_t mem_type, const T* v, hsize_t n)
    {
        Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose);
        Hid a(H5Acreate2(obj, name, mem_type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        check(a.ok(), name);
        if (!bad) check(H5Awrite(a.id, mem_type, v) >= 0, name);
    }

    // dataset [n] of `mem_type` written from `data`; returns the open dataset (closed by the caller)
    hid_t dataset(hid_t parent, const char* name, hid_t mem_type, const void* data, hsize_t n)
    {
        Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose);
        const hid_t d = H5Dcreate2(parent, name, mem_type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        check(d >= 0, name);
        if (d >= 0 && n > 0) check(H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0, name);
        return d;
    }

    void record_attrs(hid_t rec, const double ud[7])
    {
        attr_array(rec, "unitDimension", H5T_NATIVE_DOUBLE, ud, 7);
        attr_scalar(rec, "timeOffset", H5T_NATIVE_FLOAT, 0.0f);
    }

    void scalar_record(hid_t species, const char* name, hid_t mem_type, const void* data, hsize_t n,
                       const double ud[7])
    {
        Hid d(dataset(species, name, mem_type, data, n), H5Dclose);
        if (bad) return;
        record_attrs(d.id, ud);
        attr_scalar(d.id, "unitSI", H5T_NATIVE_DOUBLE, 1.0);
    }

    void vector_record(hid_t species, const char* name, hid_t mem_type, const void* const comp[3], hsize_t n,
                       const double ud[7])
    {
        Hid g(H5Gcreate2(species, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        check(g.ok(), name);
        if (bad) return;
        record_attrs(g.id, ud);
        static const char* axes[3] = {"x", "y", "z"};
        for (int a = 0; a < 3 && !bad; ++a) {
            Hid d(dataset(g.id, axes[a], mem_type, comp[a], n), H5Dclose);
            if (!bad) attr_scalar(d.id, "unitSI", H5T_NATIVE_DOUBLE, 1.0);
        }
    }

    void constant_vector_record(hid_t species, const char* name, double value, hsize_t n, const double ud[7])
    {
        Hid g(H5Gcreate2(species, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        check(g.ok(), name);
        if (bad) return;
        record_attrs(g.id, ud);
        static const char* axes[3] = {"x", "y", "z"};
        const uint64_t shape = n;
        for (int a = 0; a < 3 && !bad; ++a) {
            Hid c(H5Gcreate2(g.id, axes[a], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            check(c.ok(), axes[a]);
            if (bad) return;
            attr_scalar(c.id, "value", H5T_NATIVE_DOUBLE, value);
            attr_array(c.id, "shape", H5T_NATIVE_UINT64, &shape, 1);
            attr_scalar(c.id, "unitSI", H5T_NATIVE_DOUBLE, 1.0);
        }
    }
};

// SI unit dimensions, in the standard's order (L, M, T, I, theta, N, J)
const double kNone[7] = {0, 0, 0, 0, 0, 0, 0};
const double kLength[7] = {1, 0, 0, 0, 0, 0, 0};
const double kMass[7] = {0, 1, 0, 0, 0, 0, 0};
const double kVelocity[7] = {1, 0, -1, 0, 0, 0, 0};
const double kDensity[7] = {-3, 1, 0, 0, 0, 0, 0};
const double kSpecificEnergy[7] = {2, 0, -2, 0, 0, 0, 0};
const double kTesla[7] = {0, 1, -2, -1, 0, 0, 0};          // kg s^-2 A^-1
const double kTeslaVelocity[7] = {1, 1, -3, -1, 0, 0, 0};  // the cleaning field psi: B times a speed

std::string now_string()
{
    char buf[64];
    const std::time_t t = std::time(nullptr);
    std::tm tmv;
    localtime_r(&t, &tmv);
    std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S %z", &tmv);
    return buf;
}

template <class T>
std::vector<T> pick(const T* src, const std::vector<int64_t>& idx)
{
    std::vector<T> out(idx.size());
    for (size_t k = 0; k < idx.size(); ++k) out[k] = src[idx[k]];
    return out;
}

}  // namespace

extern "C" {

int32_t sph_openpmd_file_name(const char* pattern, uint64_t iteration, char* out, int32_t cap)
{
    if (!pattern || !out || cap < 1) return -1;
    std::string s(pattern), res;
    bool found = false;
    for (size_t i = 0; i < s.size(); ++i) {
        if (!found && s[i] == '%') {
            size_t j = i + 1;
            int width = 0;
            bool zero = false;
            if (j < s.size() && s[j] == '0') {
                zero = true;
                ++j;
            }
            while (j < s.size() && s[j] >= '0' && s[j] <= '9') width = width * 10 + (s[j++] - '0');
            if (j < s.size() && s[j] == 'T') {
                char num[64];
                std::snprintf(num, sizeof(num), zero ? "%0*llu" : "%*llu", width, (unsigned long long)iteration);
                res += num;
                i = j;
                found = true;
                continue;
            }
        }
        res += s[i];
    }
    if (!found || (int64_t)res.size() + 1 > cap) return -1;
    std::memcpy(out, res.c_str(), res.size() + 1);
    return (int32_t)res.size();
}

int sph_openpmd_write_soa(const sph_host_soa* s, int32_t mhd, const char* pattern, uint64_t iteration, double time,
                          double dt)
{
    if (!s || !pattern) return fail(SPH_E_INVALID, "null argument");
    if (s->n < 0) return fail(SPH_E_INVALID, "negative particle count");
    if (!s->id || !s->mask || !s->species || !s->pos[0] || !s->pos[1] || !s->pos[2] || !s->mass || !s->vel[0] ||
        !s->vel[1] || !s->vel[2] || !s->density || !s->internal_energy)
        return fail(SPH_E_INVALID, "id, mask, species, pos, mass, vel, density, internal_energy are required");
    if (mhd && (!s->bfield[0] || !s->bfield[1] || !s->bfield[2] || !s->psi))
        return fail(SPH_E_INVALID, "bfield and psi are required for MHD output");
    std::vector<char> fname(4096);
    if (sph_openpmd_file_name(pattern, iteration, fname.data(), (int32_t)fname.size()) < 0)
        return fail(SPH_E_INVALID, "file pattern needs a %T token (e.g. output_%06T.h5)");

    Hid file(H5Fcreate(fname.data(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(SPH_E_INVALID, std::string("cannot create ") + fname.data());
    Writer w;
    w.attr_string(file.id, "openPMD", "1.1.0");
    w.attr_scalar(file.id, "openPMDextension", H5T_NATIVE_UINT32, (uint32_t)0);
    w.attr_string(file.id, "basePath", "/data/%T/");
    w.attr_string(file.id, "particlesPath", "particles/");
    w.attr_string(file.id, "iterationEncoding", "fileBased");
    {  // the file-name pattern, without its directory (files stay valid when moved together)
        const std::string pat(pattern);
        const size_t slash = pat.find_last_of('/');
        w.attr_string(file.id, "iterationFormat", slash == std::string::npos ? pat : pat.substr(slash + 1));
    }
    w.attr_string(file.id, "author", "spearhed");
    w.attr_string(file.id, "software", "spearhed-mi355x");
    w.attr_string(file.id, "softwareVersion", "1");
    w.attr_string(file.id, "date", now_string());

    Hid data(H5Gcreate2(file.id, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    w.check(data.ok(), "data");
    const std::string it_name = std::to_string((unsigned long long)iteration);
    Hid it(H5Gcreate2(data.id, it_name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    w.check(it.ok(), "iteration");
    w.attr_scalar(it.id, "time", H5T_NATIVE_DOUBLE, time);
    w.attr_scalar(it.id, "dt", H5T_NATIVE_DOUBLE, dt);
    w.attr_scalar(it.id, "timeUnitSI", H5T_NATIVE_DOUBLE, 1.0);
    Hid parts(H5Gcreate2(it.id, "particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    w.check(parts.ok(), "particles");

    static const char* names[3] = {"fluid", "boundary", "tracer"};
    for (int sp = 0; sp < 3 && !w.bad; ++sp) {
        std::vector<int64_t> idx;
        for (int64_t i = 0; i < s->n; ++i)
            if (s->mask[i] && s->species[i] == sp) idx.push_back(i);
        if (idx.empty()) continue;
        const hsize_t n = idx.size();
        Hid g(H5Gcreate2(parts.id, names[sp], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        w.check(g.ok(), names[sp]);
        if (w.bad) break;
        const std::vector<uint64_t> id = pick(s->id, idx);
        w.scalar_record(g.id, "id", H5T_NATIVE_UINT64, id.data(), n, kNone);
        std::vector<double> pos[3];
        std::vector<float> vel[3], bf[3];
        for (int a = 0; a < 3; ++a) {
            pos[a] = pick(s->pos[a], idx);
            vel[a] = pick(s->vel[a], idx);
            if (mhd) bf[a] = pick(s->bfield[a], idx);
        }
        const void* pp[3] = {pos[0].data(), pos[1].data(), pos[2].data()};
        w.vector_record(g.id, "position", H5T_NATIVE_DOUBLE, pp, n, kLength);
        w.constant_vector_record(g.id, "positionOffset", 0.0, n, kLength);
        const std::vector<float> mass = pick(s->mass, idx);
        w.scalar_record(g.id, "mass", H5T_NATIVE_FLOAT, mass.data(), n, kMass);
        const void* vp[3] = {vel[0].data(), vel[1].data(), vel[2].data()};
        w.vector_record(g.id, "velocity", H5T_NATIVE_FLOAT, vp, n, kVelocity);
        const std::vector<float> rho = pick(s->density, idx);
        w.scalar_record(g.id, "mass_density", H5T_NATIVE_FLOAT, rho.data(), n, kDensity);
        const std::vector<float> u = pick(s->internal_energy, idx);
        w.scalar_record(g.id, "specific_internal_energy", H5T_NATIVE_FLOAT, u.data(), n, kSpecificEnergy);
        if (mhd) {
            const void* bp[3] = {bf[0].data(), bf[1].data(), bf[2].data()};
            w.vector_record(g.id, "magnetic_field", H5T_NATIVE_FLOAT, bp, n, kTesla);
            const std::vector<float> psi = pick(s->psi, idx);
            w.scalar_record(g.id, "psi", H5T_NATIVE_FLOAT, psi.data(), n, kTeslaVelocity);
        }
    }
    if (w.bad) return fail(SPH_E_DEVICE, "HDF5 write failed at " + w.why);
    if (H5Fflush(file.id, H5F_SCOPE_GLOBAL) < 0) return fail(SPH_E_DEVICE, "HDF5 flush failed");
    return SPH_OK;
}

int sph_write_openpmd(sph_ctx* ctx, int32_t run, const char* pattern, uint64_t iteration)
{
    sph_run_config cfg;
    int64_t n = 0;
    int rc = sph_get_run(ctx, run, &cfg, &n);
    if (rc) return fail(rc, "invalid context or run");
    std::vector<uint64_t> id(n);
    std::vector<uint8_t> mask(n), species(n);
    std::vector<double> pos[3];
    std::vector<float> vel[3], bf[3], mass(n), rho(n), h(n), u(n), psi(n);
    sph_host_soa s;
    std::memset(&s, 0, sizeof(s));
    s.n = n;
    s.id = id.data();
    s.mask = mask.data();
    s.species = species.data();
    for (int a = 0; a < 3; ++a) {
        pos[a].resize(n);
        vel[a].resize(n);
        s.pos[a] = pos[a].data();
        s.vel[a] = vel[a].data();
        if (cfg.mhd) {
            bf[a].resize(n);
            s.bfield[a] = bf[a].data();
        }
    }
    s.mass = mass.data();
    s.density = rho.data();
    s.smoothing_length = h.data();
    s.internal_energy = u.data();
    if (cfg.mhd) s.psi = psi.data();
    if ((rc = sph_download(ctx, run, &s))) return fail(rc, std::string("download: ") + sph_last_error(ctx));
    // the run's clock and dt: FP64 reductions of the context's current state (a query, no step)
    std::vector<sph_conserved> cons((size_t)sph_run_count(ctx));
    if ((rc = sph_reduce_conserved(ctx, cons.data()))) return fail(rc, std::string("reduce: ") + sph_last_error(ctx));
    const sph_conserved& c = cons[(size_t)run];
    return sph_openpmd_write_soa(&s, cfg.mhd ? 1 : 0, pattern, iteration, c.time, c.dt_next);
}

const char* sph_openpmd_last_error(void) { return g_err.c_str(); }

}  // extern "C"
