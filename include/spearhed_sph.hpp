// spearhed_sph.hpp -- C++ host interface of the gfx950 SPH-MHD particle step (header-only, C++17).
//
// The C ABI in spearhed_sph.h is the drop-in boundary; this header is the C++ layer a spearHED-style
// host drives it through. It keeps the reference's vocabulary:
//   * stage functors with the reference's names, called in Simulation::runOneStep order
//     (include/spearhed/control/Simulation.x.cpp:196-310): PushVelocity, UpdateVolumes,
//     CalculateNeighbours, UpdateDensity, UpdateHydroForces, EulerIntegrate;
//   * Simulation::runOneStep(step) (include/spearhed/control/Simulation.hpp:73);
//   * particle columns named after the reference attributes (particleId, multiMask, position,
//     mass, velocity, density, smoothingLength, internalEnergy, dvdt, dudt;
//     include/spearhed/param/speciesDefinition.param:100-110) plus the MHD fields.
// Errors surface as spearhed::amd::Error (a std::runtime_error), as the reference's host code
// reports failures with std::runtime_error (Simulation.x.cpp:319-327).
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spearhed_sph.h"

namespace spearhed::amd {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const sph_ctx* ctx, const char* call)
{
    if (rc == SPH_OK) return;
    std::string msg = std::string(call) + " failed (" + std::to_string(rc) + ")";
    if (ctx) {
        const char* e = sph_last_error(ctx);
        if (e && *e) msg += ": " + std::string(e);
    }
    throw Error(rc, msg);
}

// Host-side particle columns of one run (structure of arrays, owned by this object).
struct ParticleColumns {
    std::vector<uint64_t> particleId;
    std::vector<uint8_t> multiMask, species;
    std::vector<double> position[3];
    std::vector<float> mass, velocity[3], density, smoothingLength, internalEnergy;
    std::vector<float> bfield[3], psi;
    std::vector<float> dvdt[3], dudt, dbdt[3], dpsidt, divb, vsig;
    std::vector<uint32_t> nNeighbours;

    ParticleColumns() = default;
    explicit ParticleColumns(int64_t n) { resize(n); }

    int64_t size() const { return (int64_t)particleId.size(); }

    void resize(int64_t n)
    {
        particleId.assign(n, 0);
        multiMask.assign(n, 1);
        species.assign(n, SPH_SPECIES_DEFAULT);
        for (int a = 0; a < 3; ++a) {
            position[a].assign(n, 0.0);
            velocity[a].assign(n, 0.0f);
            bfield[a].assign(n, 0.0f);
            dvdt[a].assign(n, 0.0f);
            dbdt[a].assign(n, 0.0f);
        }
        for (auto* v : {&mass, &density, &smoothingLength, &internalEnergy, &psi, &dudt, &dpsidt, &divb, &vsig})
            v->assign(n, 0.0f);
        nNeighbours.assign(n, 0);
    }

    sph_host_soa view()
    {
        sph_host_soa s{};
        s.n = size();
        s.id = particleId.data();
        s.mask = multiMask.data();
        s.species = species.data();
        for (int a = 0; a < 3; ++a) {
            s.pos[a] = position[a].data();
            s.vel[a] = velocity[a].data();
            s.bfield[a] = bfield[a].data();
            s.dvdt[a] = dvdt[a].data();
            s.dbdt[a] = dbdt[a].data();
        }
        s.mass = mass.data();
        s.density = density.data();
        s.smoothing_length = smoothingLength.data();
        s.internal_energy = internalEnergy.data();
        s.psi = psi.data();
        s.dudt = dudt.data();
        s.dpsidt = dpsidt.data();
        s.divb = divb.data();
        s.vsig = vsig.data();
        s.n_neighbours = nNeighbours.data();
        return s;
    }
};

// Device-resident runs (1 = single run, > 1 = ragged batch), owning an sph_ctx.
class Context {
public:
    Context(const std::vector<sph_run_config>& runs, const std::vector<int64_t>& capacity,
            int32_t precision = SPH_PREC_MIXED, void* stream = nullptr)
        : sizes_(capacity)
    {
        if (runs.size() != capacity.size()) throw Error(SPH_E_INVALID, "runs and capacity differ in length");
        sph_ctx* h = nullptr;
        const int rc = sph_ctx_create(runs.data(), capacity.data(), (int32_t)runs.size(), precision, stream, &h);
        ctx_ = h;
        if (rc != SPH_OK) {
            std::string msg = "sph_ctx_create failed (" + std::to_string(rc) + ")";
            if (h) msg += std::string(": ") + sph_last_error(h);
            sph_ctx_destroy(h);
            ctx_ = nullptr;
            throw Error(rc, msg);
        }
    }
    ~Context()
    {
        if (ctx_) sph_ctx_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    Context(Context&& o) noexcept : ctx_(std::exchange(o.ctx_, nullptr)), sizes_(std::move(o.sizes_)) {}

    sph_ctx* handle() const { return ctx_; }
    int32_t runs() const { return (int32_t)sizes_.size(); }
    int64_t size(int32_t run) const { return sizes_.at(run); }

    void upload(int32_t run, ParticleColumns& p)
    {
        sph_host_soa s = p.view();
        check(sph_upload(ctx_, run, &s), ctx_, "sph_upload");
    }
    // Current device order (supercell-sorted): compare by particleId, never by position.
    ParticleColumns download(int32_t run)
    {
        ParticleColumns p(size(run));
        sph_host_soa s = p.view();
        check(sph_download(ctx_, run, &s), ctx_, "sph_download");
        return p;
    }
    void prime() { check(sph_prime(ctx_), ctx_, "sph_prime"); }
    void step(int32_t n = 1) { check(sph_step(ctx_, n), ctx_, "sph_step"); }
    void sync() { check(sph_sync(ctx_), ctx_, "sph_sync"); }
    void graphCapture() { check(sph_graph_capture(ctx_), ctx_, "sph_graph_capture"); }
    void graphLaunch(int32_t n) { check(sph_graph_launch(ctx_, n), ctx_, "sph_graph_launch"); }
    std::vector<double> minDt()
    {
        std::vector<double> dt(runs());
        check(sph_reduce_min_dt(ctx_, dt.data()), ctx_, "sph_reduce_min_dt");
        return dt;
    }
    std::vector<sph_conserved> conserved()
    {
        std::vector<sph_conserved> c(runs());
        check(sph_reduce_conserved(ctx_, c.data()), ctx_, "sph_reduce_conserved");
        return c;
    }
    sph_grid_info grid(int32_t run)
    {
        sph_grid_info g{};
        check(sph_get_grid(ctx_, run, &g), ctx_, "sph_get_grid");
        return g;
    }
    sph_tuning autotune(int32_t steps = 2)
    {
        sph_tuning t{};
        check(sph_autotune(ctx_, steps, &t), ctx_, "sph_autotune");
        return t;
    }

private:
    sph_ctx* ctx_ = nullptr;
    std::vector<int64_t> sizes_;
};

// ---- stage functors (reference names; each is one asynchronous call on the context's stream) ----

// r += v dt for Movable species (include/spearhed/particles/pusher/PushVelocity.hpp:35-42).
// dt <= 0 uses each run's device-resident CFL dt.
struct PushVelocity {
    void operator()(Context& c, double dt = 0.0) const { check(sph_stage_drift(c.handle(), dt), c.handle(), "sph_stage_drift"); }
};

// Supercell re-sort; stands where the region AABB refit was
// (packages/pmacc/include/spmacc/particles/regions/RegionBoundsUpdate.hpp:175-214).
struct UpdateVolumes {
    void operator()(Context& c) const { check(sph_stage_resort(c.handle()), c.handle(), "sph_stage_resort"); }
};

// Per-particle neighbour counts within H: the per-particle form of calculateNeighboursSender's
// region CSR (packages/pmacc/include/spmacc/particles/regions/NeighbourRegions.hpp:233-255).
struct CalculateNeighbours {
    std::vector<uint32_t> operator()(Context& c, int32_t run) const
    {
        check(sph_stage_neighbour_counts(c.handle()), c.handle(), "sph_stage_neighbour_counts");
        return c.download(run).nNeighbours;
    }
};

// rho_i = m_i W(0, h_i) + sum_j m_j W(r_ij, h_i) (include/spearhed/particles/density/DensitySummation.hpp:104-138).
struct UpdateDensity {
    void operator()(Context& c) const { check(sph_stage_density(c.handle()), c.handle(), "sph_stage_density"); }
};

// Symmetric pressure forces and du/dt (include/spearhed/sph/HydroForces.hpp:183-214), plus Monaghan
// viscosity, SPMHD stress and induction, and div-B cleaning when the run enables them.
struct UpdateHydroForces {
    void operator()(Context& c) const { check(sph_stage_forces(c.handle()), c.handle(), "sph_stage_forces"); }
};

// v += dvdt dt, u += dudt dt, B += dBdt dt, psi += dpsidt dt, then the FP64 CFL dt and
// conservation sums (include/spearhed/particles/pusher/EulerIntegrate.hpp:43-52).
struct EulerIntegrate {
    void operator()(Context& c, double dt = 0.0) const
    {
        check(sph_stage_integrate(c.handle(), dt), c.handle(), "sph_stage_integrate");
    }
};

// The step driver: runOneStep composes the stages in the reference's order and ends with one
// synchronisation, as Simulation::runOneStep ends with one syncWait (Simulation.x.cpp:303-307).
class Simulation {
public:
    Simulation(const std::vector<sph_run_config>& runs, const std::vector<int64_t>& capacity,
               int32_t precision = SPH_PREC_MIXED)
        : ctx_(runs, capacity, precision)
    {
    }

    Context& context() { return ctx_; }

    // Upload initial conditions and compute the first density, forces and CFL dt.
    void init(std::vector<ParticleColumns>& ics)
    {
        for (int32_t r = 0; r < ctx_.runs(); ++r) ctx_.upload(r, ics.at(r));
        ctx_.prime();
    }

    // Called after every step whose number is a multiple of `period` (the reference's plugin
    // notification, PluginConnector::setNotificationPeriod as OpenPMDPlugin.x.cpp uses it), e.g.
    // with sph_write_openpmd(ctx.handle(), run, "output_%06T.h5", step) from spearhed_openpmd.h.
    void setOutput(uint32_t period, std::function<void(Context&, uint32_t)> notify)
    {
        period_ = period;
        notify_ = std::move(notify);
    }

    // One step through the individual stage calls (what sph_step does in one call).
    void runOneStep(uint32_t step)
    {
        PushVelocity{}(ctx_);
        UpdateVolumes{}(ctx_);
        UpdateDensity{}(ctx_);
        UpdateHydroForces{}(ctx_);
        EulerIntegrate{}(ctx_);
        ctx_.sync();
        if (period_ > 0 && notify_ && step % period_ == 0) notify_(ctx_, step);
    }

    // n steps through sph_step (identical arithmetic, fewer host calls).
    void run(int32_t n)
    {
        ctx_.step(n);
        ctx_.sync();
    }

private:
    Context ctx_;
    uint32_t period_ = 0;
    std::function<void(Context&, uint32_t)> notify_;
};

}  // namespace spearhed::amd
