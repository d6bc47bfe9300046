// sph_kernels.hip -- hand-written gfx950 kernels of the SPH-MHD particle step.
//
// Stage -> kernels (reference stage in brackets):
//   drift + bin   k_drift_bin        [PushVelocity, PushVelocity.hpp:35-42; + supercell key/histogram]
//   resort        k_scan_local/top/add, k_scatter, k_gather
//                                    [UpdateVolumes RegionBoundsUpdate.hpp:175-214 and the frame index
//                                     FrameIndex.hpp:209-296, replaced by a counting sort by (cell, id)]
//   density       k_density<CH>      [UpdateDensity, DensitySummation.hpp:104-138]
//   forces        k_forces<CH, MHD>  [UpdateHydroForces, HydroForces.hpp:183-214, + AV/MHD/cleaning]
//   integrate     k_kick, k_finalize [EulerIntegrate, EulerIntegrate.hpp:43-52, + CFL and FP64 sums]
//
// Pair kernels: one wave64 per target supercell. The 27 neighbour supercells form a virtual list
// that is staged chunk by chunk into LDS (interaction records, pre-shifted by the periodic image or
// supercell-origin offset, ParticleParticleInteraction.hpp:236-264). The cell's targets are spread
// over the wave as (target t, sub-lane k) with K = 2^lk sub-lanes per target, so a supercell of
// ~14 particles still fills the wave; sub-lane k culls candidates k, k+K, ... into a 64-bit mask
// and then runs the pair physics only over set bits. The K partial sums are combined with a fixed
// xor-shuffle tree, so every result depends only on the run's own data (bit-identical whether the
// run is alone or batched with others).
#include "sph_device.hpp"
#include "sph_launch.hpp"

namespace sph {

// ------------------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int find_run_by_cell(const RunDev* runs, int nruns, int cell)
{
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (runs[mid].c0 <= cell) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// XCD-aware mapping: the hardware deals consecutive workgroups round-robin to the 8 XCDs; give
// each XCD one contiguous range of supercells so that neighbouring cells share an L2.
__device__ __forceinline__ int xcd_cell(int bid, int nblk_total, int ncell)
{
    int per = nblk_total >> 3;
    int cell = (bid & 7) * per + (bid >> 3);
    return cell < ncell ? cell : -1;
}

// ------------------------------------------------------------------------------------------------
// drift + supercell key + histogram rank
// ------------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(256) void k_drift_bin(const RunDev* __restrict__ runs, StateBuf S, WorkBuf W,
                                                   double dt_host, int do_drift)
{
    const int b = blockIdx.x;
    const int r = W.blk_run[b];
    const RunDev& R = runs[r];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    if (local >= R.np) return;
    const int64_t i = R.p0 + local;
    const double dt = dt_host > 0.0 ? dt_host : W.dt[r];
    const uint8_t live = S.mask[i];
    const bool mov = do_drift && is_movable(S.species[i], live);
    int c[3];
    for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
        int ci;
        if (MIXED) {
            double x = S.pos64[a][i];
            if (mov) {
                x = x + (double)S.vel[a][i] * dt;
                if (R.periodic[a]) {
                    double hi = R.lo[a] + R.L[a];
                    if (x >= hi) x = x - R.L[a];
                    else if (x < R.lo[a]) x = x + R.L[a];
                }
                S.pos64[a][i] = x;
            }
            ci = (int)floor((x - R.lo[a]) * R.inv_s[a]);
        } else {
            float x = S.pos32[a][i];
            if (mov) {
                x = x + S.vel[a][i] * (float)dt;
                if (R.periodic[a]) {
                    float hi = R.lo32[a] + R.L32[a];
                    if (x >= hi) x = x - R.L32[a];
                    else if (x < R.lo32[a]) x = x + R.L32[a];
                }
                S.pos32[a][i] = x;
            }
            ci = (int)floorf((x - R.lo32[a]) * R.inv_s32[a]);
        }
        ci = ci < 0 ? 0 : (ci > R.n[a] - 1 ? R.n[a] - 1 : ci);
        c[a] = ci;
    }
    const uint32_t key = (uint32_t)(R.c0 + c[0] + R.n[0] * (c[1] + R.n[1] * c[2]));
    W.key[i] = key;
    W.rank[i] = atomicAdd(&W.count[key], 1u);
}

// ------------------------------------------------------------------------------------------------
// exclusive scan of the per-cell histogram -> cell starts (absolute particle index)
// ------------------------------------------------------------------------------------------------
constexpr int kScanPerThread = 8;
constexpr int kScanBlock = 256 * kScanPerThread;

__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* __restrict__ count, uint32_t* __restrict__ cs,
                                                    uint32_t* __restrict__ bsum, int n)
{
    __shared__ uint32_t wsum[4];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int e = base + k;
        v[k] = e < n ? count[e] : 0u;
        s += v[k];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = wave_incl_scan((int)s);
    if (lane == 63) wsum[wid] = (uint32_t)incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wid; ++w) woff += wsum[w];
    uint32_t run = woff + (uint32_t)incl - s;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int e = base + k;
        if (e < n) cs[e] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = woff + (uint32_t)incl;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, int nb, uint32_t* __restrict__ cs,
                                                   int n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int e = base + threadIdx.x;
        uint32_t s = e < nb ? bsum[e] : 0u;
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        int incl = wave_incl_scan((int)s);
        if (lane == 63) wsum[wid] = (uint32_t)incl;
        __syncthreads();
        uint32_t woff = carry;
        for (int w = 0; w < wid; ++w) woff += wsum[w];
        if (e < nb) bsum[e] = woff + (uint32_t)incl - s;
        __syncthreads();
        if (threadIdx.x == 1023) carry = woff + (uint32_t)incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) cs[n] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ cs, const uint32_t* __restrict__ bsum, int n)
{
    const int base = blockIdx.x * kScanBlock;
    const uint32_t off = bsum[blockIdx.x];
    for (int k = threadIdx.x; k < kScanBlock; k += 256) {
        int e = base + k;
        if (e < n) cs[e] += off;
    }
}

// ------------------------------------------------------------------------------------------------
// scatter (histogram slot) and gather (final slot = cell start + rank of id within the cell)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(const RunDev* __restrict__ runs, WorkBuf W)
{
    const int b = blockIdx.x;
    const RunDev& R = runs[W.blk_run[b]];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    if (local >= R.np) return;
    const int64_t i = R.p0 + local;
    W.tmp[W.cs[W.key[i]] + W.rank[i]] = (uint32_t)i;
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_gather(const RunDev* __restrict__ runs, StateBuf Sold, StateBuf Snew,
                                                WorkBuf W, int has_mhd)
{
    const int b = blockIdx.x;
    const RunDev& R = runs[W.blk_run[b]];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    if (local >= R.np) return;
    const int64_t q = R.p0 + local;
    const uint32_t i = W.tmp[q];
    const uint32_t c = W.key[i];
    const uint32_t cb = W.cs[c], ce = W.cs[c + 1];
    const uint64_t myid = Sold.id[i];
    uint32_t rank = 0;
    for (uint32_t m = cb; m < ce; ++m) rank += (Sold.id[W.tmp[m]] < myid) ? 1u : 0u;
    const uint32_t f = cb + rank;
    if (rank == 0) W.count[c] = 0u;  // histogram reset for the next resort

    Snew.id[f] = myid;
    const uint8_t live = Sold.mask[i], sp = Sold.species[i];
    Snew.mask[f] = live;
    Snew.species[f] = sp;
    const float m = Sold.mass[i];
    Snew.mass[f] = m;
    Snew.h[f] = Sold.h[i];
    Snew.u[f] = Sold.u[i];
    Snew.rho[f] = Sold.rho[i];
    float pc[3];
    const int lc = (int)c - R.c0;
    const int cxyz[3] = {lc % R.n[0], (lc / R.n[0]) % R.n[1], lc / (R.n[0] * R.n[1])};
    for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
        Snew.vel[a][f] = Sold.vel[a][i];
        if (MIXED) {
            double x = Sold.pos64[a][i];
            Snew.pos64[a][f] = x;
            pc[a] = (float)((x - R.lo[a]) - (double)cxyz[a] * R.s[a]);
        } else {
            float x = Sold.pos32[a][i];
            Snew.pos32[a][f] = x;
            pc[a] = x;
        }
    }
    if (has_mhd) {
        for (int a = 0; a < 3; ++a) Snew.B[a][f] = Sold.B[a][i];
        Snew.psi[f] = Sold.psi[i];
    }
    float4 p0 = is_source(sp, live) ? make_float4(pc[0], pc[1], pc[2], m)
                                    : make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
    W.P0[f] = p0;
}

// ------------------------------------------------------------------------------------------------
// shared set-up of the 27-supercell virtual neighbour list (one wave)
// ------------------------------------------------------------------------------------------------
struct NbList {
    int start[27];
    int off[28];
    float sh[27][3];
};

template <bool MIXED>
__device__ __forceinline__ void build_nblist(const RunDev& R, int lcell, const uint32_t* __restrict__ cs, NbList& L)
{
    const int lane = threadIdx.x;
    const int c[3] = {lcell % R.n[0], (lcell / R.n[0]) % R.n[1], lcell / (R.n[0] * R.n[1])};
    int cnt = 0, st = 0;
    float sh[3] = {0.f, 0.f, 0.f};
    if (lane < 27) {
        const int d[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
        int idx[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            int x = c[a] + d[a];
            int wrap = 0;
            if (R.periodic[a]) {
                wrap = x < 0 ? -1 : (x >= R.n[a] ? 1 : 0);
                x -= wrap * R.n[a];
            } else if (x < 0 || x >= R.n[a]) {
                ok = false;
            }
            idx[a] = x;
            sh[a] = MIXED ? (float)d[a] * R.s32[a] : (float)wrap * R.L32[a];
        }
        if (ok) {
            int cc = R.c0 + idx[0] + R.n[0] * (idx[1] + R.n[1] * idx[2]);
            st = (int)cs[cc];
            cnt = (int)cs[cc + 1] - st;
        }
    }
    int incl = wave_incl_scan(cnt);
    if (lane < 27) {
        L.start[lane] = st;
        L.off[lane] = incl - cnt;
        L.sh[lane][0] = sh[0];
        L.sh[lane][1] = sh[1];
        L.sh[lane][2] = sh[2];
    }
    if (lane == 26) L.off[27] = incl;
    __syncthreads();
}

// segment of virtual index v (largest k with off[k] <= v, over non-decreasing off[0..26])
__device__ __forceinline__ int nb_segment(const NbList& L, int v)
{
    int k = 0;
#pragma unroll
    for (int step = 16; step >= 1; step >>= 1) {
        int t = k + step;
        if (t < 27 && L.off[t] <= v) k = t;
    }
    return k;
}

__device__ __forceinline__ int choose_lk(int nt, int kmax_log)
{
    // K = 64 / next_pow2(nt), capped at 2^kmax_log; K = 1 for nt >= 64
    int np2 = 1, lk = 6;
    while (np2 < nt && lk > 0) {
        np2 <<= 1;
        --lk;
    }
    return lk < kmax_log ? lk : kmax_log;
}

// ------------------------------------------------------------------------------------------------
// density pass (+ per-particle interaction records for the force pass)
// ------------------------------------------------------------------------------------------------
template <int CH, bool MIXED, bool MHD>
__global__ __launch_bounds__(64) void k_density(const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                                int nblk_total, StateBuf S, WorkBuf W, int kmax_log)
{
    __shared__ NbList L;
    __shared__ float4 sp[CH];
    const int cell = xcd_cell(blockIdx.x, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    build_nblist<MIXED>(R, cell - R.c0, W.cs, L);
    const int total = L.off[27];
    const int self_v0 = L.off[13];
    const int lane = threadIdx.x;
    const int lk = choose_lk(nt, kmax_log);
    const int K = 1 << lk;
    const int tpg = 64 >> lk;
    const int kern = R.kernel;
    const float H2 = R.H2;
    const float gam = R.gamma;

    for (int tg = 0; tg < nt; tg += tpg) {
        const int t = tg + (lane >> lk);
        const int k = lane & (K - 1);
        const bool valid = t < nt;
        const int i = t0 + (valid ? t : 0);
        const uint8_t live = S.mask[i], spc = S.species[i];
        const bool tgt = valid && is_target(spc, live);
        const float4 own = W.P0[i];
        const float h = S.h[i];
        const float invh = 1.0f / h;
        const int self_v = self_v0 + t;
        float acc = 0.0f;
        uint32_t cnt = 0;

        for (int base = 0; base < total; base += CH) {
            for (int e = lane; e < CH; e += 64) {
                int v = base + e;
                float4 p = make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
                if (v < total) {
                    int sg = nb_segment(L, v);
                    p = W.P0[L.start[sg] + (v - L.off[sg])];
                    p.x += L.sh[sg][0];
                    p.y += L.sh[sg][1];
                    p.z += L.sh[sg][2];
                }
                sp[e] = p;
            }
            __syncthreads();
            const int per = CH >> lk;
            for (int sb = 0; sb < per; sb += 64) {
                const int lim = per - sb < 64 ? per - sb : 64;
                uint64_t msk = 0;
                for (int ii = 0; ii < lim; ++ii) {
                    const int e = k + ((sb + ii) << lk);
                    float dx, dy, dz;
                    const float r2 = geom_r2(own.x, own.y, own.z, sp[e], dx, dy, dz);
                    const bool ok = (r2 < H2) && (base + e != self_v);
                    msk |= (uint64_t)ok << ii;
                }
                if (!tgt) msk = 0;
                cnt += (uint32_t)__popcll(msk);
                while (msk) {
                    const int ii = __builtin_ctzll(msk);
                    msk &= msk - 1;
                    const int e = k + ((sb + ii) << lk);
                    const float4 p = sp[e];
                    float dx, dy, dz;
                    const float r2 = geom_r2(own.x, own.y, own.z, p, dx, dy, dz);
                    const float invR = r2 > 0.0f ? fast_rsq(r2) : 0.0f;
                    const float rr = r2 * invR;
                    acc += p.w * kshape(kern, rr * invh);
                }
            }
            __syncthreads();
        }
        for (int off = 1; off < K; off <<= 1) {
            acc += __shfl_xor(acc, off, 64);
            cnt += __shfl_xor(cnt, off, 64);
        }
        // epilogue: one lane per particle writes rho and its interaction records
        float cmax = 0.0f;
        if (k == 0 && valid && is_source(spc, live)) {
#pragma clang fp contract(off)
            const float sig = R.sigma;
            const float wfac = sig / (h * h * h);
            const float gfac = (kern == 2) ? -5.0f * sig / (h * h * h * h * h) : sig / (h * h * h * h);
            float rho;
            if (tgt) {
                rho = wfac * (acc + own.w * R.f0);
                S.rho[i] = rho;
                W.nnb[i] = cnt;
            } else {
                rho = S.rho[i];
            }
            const float uu = S.u[i];
            const float P = (gam - 1.0f) * rho * uu;
            const float invrho = 1.0f / rho;
            float B2 = 0.0f, bx = 0.f, by = 0.f, bz = 0.f, psi = 0.f;
            if (MHD && R.mhd) {
                bx = S.B[0][i];
                by = S.B[1][i];
                bz = S.B[2][i];
                psi = S.psi[i];
                B2 = bx * bx + by * by + bz * bz;
            }
            const float bdr = invrho * invrho;
            const float ph = P * bdr;
            const float pr = (P + 0.5f * B2) * bdr;
            const float psr = psi * bdr;
            const float c = sqrtf((gam * P + B2) * invrho);
            W.P1[i] = make_float4(S.vel[0][i], S.vel[1][i], S.vel[2][i], c);
            W.P2[i] = make_float4(h, invh, gfac, rho);
            W.P3[i] = make_float4(pr, ph, bdr, psr);
            if (MHD) W.P4[i] = make_float4(bx, by, bz, psi);
            if (tgt) cmax = c;
        }
        if (k == 0 && valid && !tgt) W.nnb[i] = 0;
        if (MHD && R.mhd && R.cleaning) {
            for (int off = 32; off >= 1; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off, 64));
            if (lane == 0 && cmax > 0.0f) atomicMax(&W.chbits[r], __float_as_uint(cmax));
        }
    }
}

// ------------------------------------------------------------------------------------------------
// force pass: pressure gradient, P dV, Monaghan viscosity, SPMHD stress, induction, cleaning
// ------------------------------------------------------------------------------------------------
struct ForceAcc {
    float ax, ay, az, DV, AVU, vmax;
    float Ti, Tj, TBx, TBy, TBz, VBx, VBy, VBz, DB, PSx, PSy, PSz;
};

template <int CH, bool MIXED, bool MHD>
__global__ __launch_bounds__(64) void k_forces(const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                               int nblk_total, StateBuf S, WorkBuf W, int kmax_log)
{
    constexpr int NP = MHD ? 5 : 4;
    __shared__ NbList L;
    __shared__ float4 sp[NP][CH];
    const int cell = xcd_cell(blockIdx.x, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    build_nblist<MIXED>(R, cell - R.c0, W.cs, L);
    const int total = L.off[27];
    const int self_v0 = L.off[13];
    const int lane = threadIdx.x;
    const int lk = choose_lk(nt, kmax_log);
    const int K = 1 << lk;
    const int tpg = 64 >> lk;
    const int kern = R.kernel;
    const float H2 = R.H2;
    const float alpha = R.alpha, beta = R.beta;
    const bool mhd = MHD && R.mhd;
    const bool clean = mhd && R.cleaning;
    const float borve = mhd ? R.borve : 0.0f;
    const float ch = __uint_as_float(W.chbits[r]);

    for (int tg = 0; tg < nt; tg += tpg) {
        const int t = tg + (lane >> lk);
        const int k = lane & (K - 1);
        const bool valid = t < nt;
        const int i = t0 + (valid ? t : 0);
        const bool tgt = valid && is_target(S.species[i], S.mask[i]);
        const float4 o0 = W.P0[i], o1 = W.P1[i], o2 = W.P2[i], o3 = W.P3[i];
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MHD) o4 = W.P4[i];
        const int self_v = self_v0 + t;
        ForceAcc A = {};

        for (int base = 0; base < total; base += CH) {
            for (int e = lane; e < CH; e += 64) {
                int v = base + e;
                if (v < total) {
                    int sg = nb_segment(L, v);
                    int j = L.start[sg] + (v - L.off[sg]);
                    float4 p = W.P0[j];
                    p.x += L.sh[sg][0];
                    p.y += L.sh[sg][1];
                    p.z += L.sh[sg][2];
                    sp[0][e] = p;
                    sp[1][e] = W.P1[j];
                    sp[2][e] = W.P2[j];
                    sp[3][e] = W.P3[j];
                    if (MHD) sp[4][e] = W.P4[j];
                } else {
                    sp[0][e] = make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
                }
            }
            __syncthreads();
            const int per = CH >> lk;
            for (int sb = 0; sb < per; sb += 64) {
                const int lim = per - sb < 64 ? per - sb : 64;
                uint64_t msk = 0;
                for (int ii = 0; ii < lim; ++ii) {
                    const int e = k + ((sb + ii) << lk);
                    float dx, dy, dz;
                    const float r2 = geom_r2(o0.x, o0.y, o0.z, sp[0][e], dx, dy, dz);
                    const bool ok = (r2 < H2) && (base + e != self_v);
                    msk |= (uint64_t)ok << ii;
                }
                if (!tgt) msk = 0;
                while (msk) {
                    const int ii = __builtin_ctzll(msk);
                    msk &= msk - 1;
                    const int e = k + ((sb + ii) << lk);
                    const float4 q0 = sp[0][e], q1 = sp[1][e], q2 = sp[2][e], q3 = sp[3][e];
                    float dx, dy, dz;
                    const float r2 = geom_r2(o0.x, o0.y, o0.z, q0, dx, dy, dz);
                    const float invR = r2 > 0.0f ? fast_rsq(r2) : 0.0f;
                    const float rr = r2 * invR;
                    const float s_i = kgrad(kern, rr, invR, rr * o2.y, o2.z);
                    const float s_j = kgrad(kern, rr, invR, rr * q2.y, q2.z);
                    const float dvx = o1.x - q1.x, dvy = o1.y - q1.y, dvz = o1.z - q1.z;
                    const float vr = dvx * dx + dvy * dy + dvz * dz;
                    const float m = q0.w;
                    const float ms_i = m * s_i, ms_j = m * s_j;
                    const float fiso = o3.x * ms_i + q3.x * ms_j;
                    A.ax -= fiso * dx;
                    A.ay -= fiso * dy;
                    A.az -= fiso * dz;
                    A.DV += ms_i * vr;
                    const float w = vr * invR;
                    const float vs = o1.w + q1.w - beta * fminf(w, 0.0f);
                    A.vmax = fmaxf(A.vmax, vs);
                    if (alpha > 0.0f && vr < 0.0f) {
                        const float hb = 0.5f * (o2.x + q2.x);
                        const float mu = hb * vr * fast_rcp(r2 + 0.01f * hb * hb);
                        const float cb = 0.5f * (o1.w + q1.w);
                        const float rb = 0.5f * (o2.w + q2.w);
                        const float Pi = (beta * mu - alpha * cb) * mu * fast_rcp(rb);
                        const float gv = Pi * 0.5f * (ms_i + ms_j);
                        A.ax -= gv * dx;
                        A.ay -= gv * dy;
                        A.az -= gv * dz;
                        A.AVU += gv * vr;
                    }
                    if (MHD && mhd) {
                        const float4 q4 = sp[4][e];
                        const float Bir = o4.x * dx + o4.y * dy + o4.z * dz;
                        const float Bjr = q4.x * dx + q4.y * dy + q4.z * dz;
                        const float ti = Bir * o3.z * ms_i;
                        const float tj = Bjr * q3.z * ms_j;
                        A.Ti += ti;
                        A.Tj += tj;
                        A.TBx += tj * q4.x;
                        A.TBy += tj * q4.y;
                        A.TBz += tj * q4.z;
                        const float wv = ms_i * Bir;
                        A.VBx += wv * dvx;
                        A.VBy += wv * dvy;
                        A.VBz += wv * dvz;
                        A.DB += ms_i * (Bir - Bjr);
                        if (clean) {
                            const float fps = o3.w * ms_i + q3.w * ms_j;
                            A.PSx += fps * dx;
                            A.PSy += fps * dy;
                            A.PSz += fps * dz;
                        }
                    }
                }
            }
            __syncthreads();
        }
        for (int off = 1; off < K; off <<= 1) {
            A.ax += __shfl_xor(A.ax, off, 64);
            A.ay += __shfl_xor(A.ay, off, 64);
            A.az += __shfl_xor(A.az, off, 64);
            A.DV += __shfl_xor(A.DV, off, 64);
            A.AVU += __shfl_xor(A.AVU, off, 64);
            A.vmax = fmaxf(A.vmax, __shfl_xor(A.vmax, off, 64));
            if (MHD) {
                A.Ti += __shfl_xor(A.Ti, off, 64);
                A.Tj += __shfl_xor(A.Tj, off, 64);
                A.TBx += __shfl_xor(A.TBx, off, 64);
                A.TBy += __shfl_xor(A.TBy, off, 64);
                A.TBz += __shfl_xor(A.TBz, off, 64);
                A.VBx += __shfl_xor(A.VBx, off, 64);
                A.VBy += __shfl_xor(A.VBy, off, 64);
                A.VBz += __shfl_xor(A.VBz, off, 64);
                A.DB += __shfl_xor(A.DB, off, 64);
                A.PSx += __shfl_xor(A.PSx, off, 64);
                A.PSy += __shfl_xor(A.PSy, off, 64);
                A.PSz += __shfl_xor(A.PSz, off, 64);
            }
        }
        if (k == 0 && tgt) {
            const float rho = o2.w;
            const float invrho = 1.0f / rho;
            float ax = A.ax, ay = A.ay, az = A.az;
            if (MHD && mhd) {
                const float tcoef = A.Ti - borve * (A.Ti + A.Tj);
                ax = ax + A.TBx + o4.x * tcoef;
                ay = ay + A.TBy + o4.y * tcoef;
                az = az + A.TBz + o4.z * tcoef;
            }
            W.dvdt[0][i] = ax;
            W.dvdt[1][i] = ay;
            W.dvdt[2][i] = az;
            W.dudt[i] = o3.y * A.DV + 0.5f * A.AVU;
            W.vsig[i] = fmaxf(2.0f * o1.w, A.vmax);
            if (MHD && mhd) {
                const float divB = -A.DB * invrho;
                const float divv = -A.DV * invrho;
                float dbx = -(A.VBx - o4.x * A.DV) * invrho;
                float dby = -(A.VBy - o4.y * A.DV) * invrho;
                float dbz = -(A.VBz - o4.z * A.DV) * invrho;
                float dpsi = 0.0f;
                if (clean) {
                    dbx -= rho * A.PSx;
                    dby -= rho * A.PSy;
                    dbz -= rho * A.PSz;
                    const float psi = o4.w;
                    dpsi = -ch * ch * divB - psi * R.clean_sigma * ch * o2.y - 0.5f * psi * divv;
                }
                W.dbdt[0][i] = dbx;
                W.dbdt[1][i] = dby;
                W.dbdt[2][i] = dbz;
                W.dpsidt[i] = dpsi;
                W.divb[i] = divB;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// kick + CFL + FP64 conservation sums (deterministic two-level reduction)
// ------------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(256) void k_kick(const RunDev* __restrict__ runs, StateBuf S, WorkBuf W, double dt_host,
                                              int do_kick)
{
    __shared__ double red[6][4];
    const int b = blockIdx.x;
    const int r = W.blk_run[b];
    const RunDev& R = runs[r];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    double v_dt = INFINITY, v_m = 0, v_px = 0, v_py = 0, v_pz = 0, v_e = 0;
    if (local < R.np) {
        const int64_t i = R.p0 + local;
        if (is_target(S.species[i], S.mask[i])) {
            const double dt = dt_host > 0.0 ? dt_host : W.dt[r];
            const float dtf = (float)dt;
            float vx = S.vel[0][i], vy = S.vel[1][i], vz = S.vel[2][i];
            float u = S.u[i];
            const float ax = W.dvdt[0][i], ay = W.dvdt[1][i], az = W.dvdt[2][i];
            float bx = 0.f, by = 0.f, bz = 0.f;
            if (do_kick) {
                vx = vx + ax * dtf;
                vy = vy + ay * dtf;
                vz = vz + az * dtf;
                u = u + W.dudt[i] * dtf;
                S.vel[0][i] = vx;
                S.vel[1][i] = vy;
                S.vel[2][i] = vz;
                S.u[i] = u;
            }
            if (R.mhd) {
                bx = S.B[0][i];
                by = S.B[1][i];
                bz = S.B[2][i];
                if (do_kick) {
                    bx = bx + W.dbdt[0][i] * dtf;
                    by = by + W.dbdt[1][i] * dtf;
                    bz = bz + W.dbdt[2][i] * dtf;
                    S.B[0][i] = bx;
                    S.B[1][i] = by;
                    S.B[2][i] = bz;
                    S.psi[i] = S.psi[i] + W.dpsidt[i] * dtf;
                }
            }
            const float h = S.h[i], vsg = W.vsig[i];
            if (MIXED) {
                double dti = (double)R.cfl * (double)h / (double)vsg;
                double a2 = (double)ax * ax + (double)ay * ay + (double)az * az;
                if (a2 > 0.0) dti = fmin(dti, (double)R.cfl_force * sqrt((double)h / sqrt(a2)));
                v_dt = dti;
            } else {
                float dti = R.cfl * h / vsg;
                float a2 = ax * ax + ay * ay + az * az;
                if (a2 > 0.0f) dti = fminf(dti, R.cfl_force * sqrtf(h / sqrtf(a2)));
                v_dt = (double)dti;
            }
            const double m = (double)S.mass[i];
            const double dvx = vx, dvy = vy, dvz = vz;
            const double b2 = (double)bx * bx + (double)by * by + (double)bz * bz;
            v_m = m;
            v_px = m * dvx;
            v_py = m * dvy;
            v_pz = m * dvz;
            v_e = m * (0.5 * (dvx * dvx + dvy * dvy + dvz * dvz) + (double)u + 0.5 * b2 / (double)S.rho[i]);
        }
    }
    // wave then block reduction, fixed order
    for (int off = 32; off >= 1; off >>= 1) {
        v_dt = fmin(v_dt, __shfl_xor(v_dt, off, 64));
        v_m += __shfl_xor(v_m, off, 64);
        v_px += __shfl_xor(v_px, off, 64);
        v_py += __shfl_xor(v_py, off, 64);
        v_pz += __shfl_xor(v_pz, off, 64);
        v_e += __shfl_xor(v_e, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = v_dt;
        red[1][wid] = v_m;
        red[2][wid] = v_px;
        red[3][wid] = v_py;
        red[4][wid] = v_pz;
        red[5][wid] = v_e;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        double s;
        if (q == 0) s = fmin(fmin(red[0][0], red[0][1]), fmin(red[0][2], red[0][3]));
        else s = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
        W.partial[(int64_t)b * 6 + q] = s;
    }
}

// One block per run: reduce the run's kick-block partials in a fixed tree order, publish the
// conserved sums, set the next dt, advance the time and clear the cleaning-speed maximum.
__global__ __launch_bounds__(256) void k_finalize(const RunDev* __restrict__ runs, WorkBuf W, double dt_host,
                                                  int advance)
{
    __shared__ double red[6][256];
    const int r = blockIdx.x;
    const RunDev& R = runs[r];
    double s[6] = {INFINITY, 0, 0, 0, 0, 0};
    for (int bb = threadIdx.x; bb < R.kb_count; bb += 256) {
        const double* p = W.partial + (int64_t)(R.kb_first + bb) * 6;
        s[0] = fmin(s[0], p[0]);
        for (int q = 1; q < 6; ++q) s[q] += p[q];
    }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + w]);
            for (int q = 1; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dt_used = dt_host > 0.0 ? dt_host : W.dt[r];
        double t = W.time[r];
        if (advance) t += dt_used;
        W.time[r] = t;
        const double dtn = R.dt_fixed > 0.0 ? R.dt_fixed : red[0][0];
        double* c = W.cons + (int64_t)r * 8;
        c[0] = red[1][0];
        c[1] = red[2][0];
        c[2] = red[3][0];
        c[3] = red[4][0];
        c[4] = red[5][0];
        c[5] = dtn;
        c[6] = t;
        c[7] = (double)__uint_as_float(W.chbits[r]);
        W.dt[r] = dtn;
        W.chbits[r] = 0u;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline int nblk_xcd(int ncell) { return ((ncell + 7) / 8) * 8; }

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_drift_bin(const LaunchCtx& c, double dt_host, int do_drift)
{
    if (c.mixed)
        k_drift_bin<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_drift);
    else
        k_drift_bin<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_drift);
    return hipGetLastError();
}

hipError_t launch_scan(const LaunchCtx& c)
{
    const int n = c.ncell_tot;
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    k_scan_local<<<nb, 256, 0, c.stream>>>(c.W.count, c.W.cs, c.W.bsum, n);
    k_scan_top<<<1, 1024, 0, c.stream>>>(c.W.bsum, nb, c.W.cs, n);
    k_scan_add<<<nb, 256, 0, c.stream>>>(c.W.cs, c.W.bsum, n);
    return hipGetLastError();
}

hipError_t launch_scatter_gather(const LaunchCtx& c, const StateBuf& Snew)
{
    k_scatter<<<c.nkb, 256, 0, c.stream>>>(c.runs, c.W);
    if (c.mixed)
        k_gather<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, Snew, c.W, c.has_mhd);
    else
        k_gather<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, Snew, c.W, c.has_mhd);
    return hipGetLastError();
}

template <int CH>
static void density_ch(const LaunchCtx& c, int nblk)
{
    if (c.mixed) {
        if (c.has_mhd)
            k_density<CH, true, true><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
        else
            k_density<CH, true, false><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
    } else {
        if (c.has_mhd)
            k_density<CH, false, true><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
        else
            k_density<CH, false, false><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
    }
}

hipError_t launch_density(const LaunchCtx& c)
{
    const int nblk = nblk_xcd(c.ncell_tot);
    switch (c.density_chunk) {
    case 128: density_ch<128>(c, nblk); break;
    case 512: density_ch<512>(c, nblk); break;
    default: density_ch<256>(c, nblk); break;
    }
    return hipGetLastError();
}

template <int CH>
static void forces_ch(const LaunchCtx& c, int nblk)
{
    if (c.mixed) {
        if (c.has_mhd)
            k_forces<CH, true, true><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
        else
            k_forces<CH, true, false><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
    } else {
        if (c.has_mhd)
            k_forces<CH, false, true><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
        else
            k_forces<CH, false, false><<<nblk, 64, 0, c.stream>>>(c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax_log);
    }
}

hipError_t launch_forces(const LaunchCtx& c)
{
    const int nblk = nblk_xcd(c.ncell_tot);
    switch (c.force_chunk) {
    case 64: forces_ch<64>(c, nblk); break;
    case 256: forces_ch<256>(c, nblk); break;
    default: forces_ch<128>(c, nblk); break;
    }
    return hipGetLastError();
}

hipError_t launch_kick(const LaunchCtx& c, double dt_host, int do_kick)
{
    if (c.mixed)
        k_kick<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_kick);
    else
        k_kick<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_kick);
    k_finalize<<<c.nruns, 256, 0, c.stream>>>(c.runs, c.W, dt_host, do_kick);
    return hipGetLastError();
}

}  // namespace sph
