// C++ host driving the C ABI through include/spearhed_sph.hpp, as a spearHED-style host would.
//
//   host_stepper cpu  -- checks that need no GPU: invalid configurations raise spearhed::amd::Error
//                        with SPH_E_INVALID before any device call; sizing helpers answer.
//   host_stepper gpu  -- a seeded 10^3 lattice run stepped three times through the stage functors
//                        (Simulation::runOneStep) and three times through sph_step; the two final
//                        states must be bit-identical when matched by particleId.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>

#include "spearhed_sph.hpp"

using namespace spearhed::amd;

static sph_run_config lattice_config(int n)
{
    sph_run_config c{};
    for (int a = 0; a < 3; ++a) {
        c.box_lo[a] = 0.0;
        c.box_len[a] = 1.0;
        c.periodic[a] = 1;
    }
    c.kernel = SPH_KERNEL_WENDLAND;
    c.h0 = (float)(1.2 / n);
    c.gamma = 5.0f / 3.0f;
    c.alpha_visc = 1.0f;
    c.beta_visc = 2.0f;
    c.mhd = 1;
    c.cleaning = 1;
    c.borve = 1;
    c.clean_sigma = 1.0f;
    c.cfl = 0.3f;
    c.cfl_force = 0.25f;
    c.dt_fixed = 0.0;
    c.cell_factor = 1.0f;
    return c;
}

static ParticleColumns lattice_particles(int n)
{
    const int64_t N = (int64_t)n * n * n;
    ParticleColumns p(N);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {  // xorshift64*: uniform in [0, 1)
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return (double)((s * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
    };
    const double dx = 1.0 / n;
    int64_t i = 0;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x, ++i) {
                p.particleId[i] = (uint64_t)(N - 1 - i);  // ids not in storage order
                const int g[3] = {x, y, z};
                for (int a = 0; a < 3; ++a) {
                    p.position[a][i] = (g[a] + 0.5 + 0.1 * (rnd() - 0.5)) * dx;
                    p.velocity[a][i] = (float)(0.3 * (rnd() - 0.5));
                    p.bfield[a][i] = (float)(a == 2 ? 0.1 : 0.02 * (rnd() - 0.5));
                }
                p.mass[i] = (float)(dx * dx * dx);
                p.density[i] = 1.0f;
                p.smoothingLength[i] = (float)(1.2 * dx);
                p.internalEnergy[i] = 1.5f;
            }
    return p;
}

static int cpu_checks()
{
    sph_run_config bad = lattice_config(8);
    bad.h0 = 0.0f;
    try {
        Context ctx({bad}, {512});
        std::fprintf(stderr, "expected an error for h0 = 0\n");
        return 1;
    } catch (const Error& e) {
        if (e.code() != SPH_E_INVALID || std::strlen(e.what()) == 0) {
            std::fprintf(stderr, "wrong error: %d %s\n", e.code(), e.what());
            return 1;
        }
    }
    const sph_run_config good = lattice_config(8);
    if (sph_bytes_per_run(&good, 512, SPH_PREC_MIXED) <= 0 || sph_run_cells(&good) <= 0) return 1;
    if (sph_pool_bytes(1 << 20, 1 << 16, 8, SPH_PREC_MIXED, 1) <= sph_pool_bytes(1 << 20, 1 << 16, 8, SPH_PREC_FP32, 0))
        return 1;
    std::printf("{\"cpu\": \"ok\"}\n");
    return 0;
}

static std::map<uint64_t, int64_t> index_by_id(const ParticleColumns& p)
{
    std::map<uint64_t, int64_t> m;
    for (int64_t i = 0; i < p.size(); ++i) m[p.particleId[i]] = i;
    return m;
}

static int gpu_checks()
{
    const int n = 10;
    const sph_run_config cfg = lattice_config(n);
    std::vector<ParticleColumns> ic1{lattice_particles(n)}, ic2{lattice_particles(n)};
    const int64_t N = ic1[0].size();
    Simulation staged({cfg}, {N}), fused({cfg}, {N});
    staged.init(ic1);
    fused.init(ic2);
    // the output hook fires after steps 0 and 2 (period 2), with the context the step ran on
    std::vector<uint32_t> notified;
    staged.setOutput(2, [&](Context& ctx, uint32_t step) {
        if (&ctx == &staged.context()) notified.push_back(step);
    });
    for (uint32_t s = 0; s < 3; ++s) staged.runOneStep(s);
    if (notified != std::vector<uint32_t>{0, 2}) {
        std::printf("{\"error\": \"output hook fired at the wrong steps\"}\n");
        return 1;
    }
    fused.run(3);
    ParticleColumns a = staged.context().download(0), b = fused.context().download(0);
    auto ia = index_by_id(a), ib = index_by_id(b);
    int64_t diff = 0;
    for (const auto& [id, i] : ia) {
        const int64_t j = ib.at(id);
        for (int d = 0; d < 3; ++d) {
            diff += a.position[d][i] != b.position[d][j];
            diff += std::memcmp(&a.velocity[d][i], &b.velocity[d][j], 4) != 0;
            diff += std::memcmp(&a.bfield[d][i], &b.bfield[d][j], 4) != 0;
        }
        diff += std::memcmp(&a.internalEnergy[i], &b.internalEnergy[j], 4) != 0;
        diff += std::memcmp(&a.density[i], &b.density[j], 4) != 0;
    }
    const sph_conserved ca = staged.context().conserved()[0], cb = fused.context().conserved()[0];
    std::printf("{\"particles\": %lld, \"differing_values\": %lld, \"time\": [%.17g, %.17g], \"energy\": [%.17g, %.17g]}\n",
                (long long)N, (long long)diff, ca.time, cb.time, ca.energy, cb.energy);
    return (diff == 0 && ca.time == cb.time && ca.energy == cb.energy && ca.time > 0.0) ? 0 : 1;
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    try {
        return gpu ? gpu_checks() : cpu_checks();
    } catch (const Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
