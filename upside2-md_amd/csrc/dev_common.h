// Device helpers shared by the engine kernels: block reductions, the B-spline
// and sigmoid primitives of the reference (src/spline.h, src/vector_math.h),
// and the fixed-order gather that replaces scatter-with-atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "engine_dev.h"

namespace upside {

// Execution context of one workgroup.  A system is evaluated by a team of G
// workgroups (G = 1: a single workgroup, no grid barriers).  Data-parallel
// loops stride over the team's threads (gtid, gsize); team_sync separates
// phases whose data cross workgroups.
// Address-space-typed pointers.  Arena and table pointers reach the device
// code through descriptors, so the compiler sees them as generic (flat)
// pointers: every flat load is counted against both the vector-memory and
// the LDS counters and waits serialise.  Hot loops re-type their pointers
// with gmem()/lmem() so that loads become global_load / ds_read.
#define UPS_GLOBAL __attribute__((address_space(1)))
#define UPS_LDS __attribute__((address_space(3)))
template <class T> using gp = UPS_GLOBAL T*;
template <class T> using lp = UPS_LDS T*;
template <class T> __device__ __forceinline__ gp<T> gmem(T* p) { return (gp<T>)p; }
template <class T> __device__ __forceinline__ lp<T> lmem(T* p) { return (lp<T>)p; }

struct Ctx {
    float* A;          // per-system float arena
    int* I;            // per-system int arena
    const float* CF;   // constant float arena
    const int* CI;     // constant int arena
    float* red;        // LDS reduction scratch [NWAVE]
    int* ired;         // LDS reduction scratch [NWAVE] (ints)
    float* lds;        // LDS staging area [LDS_FLOATS]
    int tid;           // thread in the workgroup
    unsigned long long* prof;  // wall-clock accumulators (null: profiling off)
    int pbase;                 // first fixed profile slot
    int cur_op;                // program op being run (team barriers are counted per op while profiling)
    int err;                   // int arena offset of the error word this (sub-)team sets and reads
    int early_done;            // this evaluation ran the early deriv ops beside the rotamer solve
    int rot_graph_done;        // this evaluation refreshed the rotamer pair graph in a section lane
    int rank, G;       // workgroup rank in the team, team size
    int gtid, gsize;   // thread in the team, team threads
    unsigned long long* bar;  // team barrier arrival counter
    unsigned long long bar_next;  // count that opens this workgroup's next barrier on `bar` (thread 0)
    unsigned* bar_flag;       // team barrier spin-timeout flag
    int local;                // the whole team runs on one XCD (shares one L2)
    unsigned long long* bar2; // arrival counter of the sub-team of ranks 1..G-1
    float* tred2;             // the sub-team's reduction scratch
    int tpar2;                // the sub-team's next reduction buffer parity
    unsigned long long* bar_lo;  // sub-team of ranks 0..G/2-1: counter, scratch, parity
    float* tred_lo;
    int tpar_lo;
    unsigned long long* bar_hi;  // sub-team of ranks G/2..G-1
    float* tred_hi;
    int tpar_hi;
    unsigned long long* bar_q;   // the four quarter sub-teams: counters bar_q[j], scratch tred_q + 2*MAX_TEAM*j,
    float* tred_q;               // and the parity of this workgroup's own quarter
    int tpar_q;
    float* tred;       // team reduction scratch, 2 buffers of G floats
    int tpar;          // buffer parity of the next team reduction
    int bp_layout;     // where rotamer solves run (see bp_on_one_wg)
    const RunArgs* args;  // the launch's arguments (rotamer solve schedule, warm-start network)
};

__device__ __forceinline__ unsigned long long ptick(const Ctx& c) { return c.prof ? wall_clock64() : 0ull; }
__device__ __forceinline__ void pacc(const Ctx& c, int slot, unsigned long long t0) {
    // a no-return atomic: the thread does not wait for the accumulator's line
    if (c.prof && c.gtid == 0)
        __hip_atomic_fetch_add(c.prof + c.pbase + slot, wall_clock64() - t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Team barrier on a monotonic 64-bit arrival counter: arrival n of a team of
// G falls in barrier n / G, which opens when the counter reaches the next
// multiple of G.  Nothing is reset between barriers, and the counter is a
// multiple of G between launches (the host zeroes it when the team size
// changes).  Every workgroup keeps the count that opens its next barrier
// (bar_base at the start, then + G per barrier), so its arrival is a
// no-return atomic: it does not wait for the counter's old value before it
// starts polling.  A stale poll can only prolong a wait.
//   Memory ordering: a team spread over XCDs (separate L2s) needs
// device-scope fences (write back L2 before arriving, invalidate after the
// wait).  A team on one XCD shares its L2: every wave waits for its stores to
// be acknowledged by L2 before the workgroup barrier (no L2 write-back), and
// after the wait a device-scope invalidate drops the CU's L1.  (A
// workgroup-scope invalidate does not reach L1 outside thread-group-split
// mode.)
//   The wait is bounded: after ~1 s of spinning the workgroup records a
// timeout and proceeds, so a launch always drains; the host then reports the
// run as failed.
constexpr unsigned BAR_SPIN_LIMIT = 1u << 24;

// s_waitcnt immediate: vmcnt(0), expcnt and lgkmcnt not waited on.
constexpr int WAIT_VMEM = 0x0F70;

// Count that opens the first barrier of a team of G on counter `bar`, read
// before this workgroup's first arrival: fewer than G arrivals of that
// barrier can precede it (it cannot open without this workgroup), so the
// count rounded down to a multiple of G is the base.  Thread 0.
__device__ __forceinline__ unsigned long long bar_first(unsigned long long* bar, int G) {
    const unsigned long long v = __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (v / (unsigned long long)G + 1ull) * (unsigned long long)G;
}

__device__ void team_sync(Ctx& c) {
    // every wave's stores acknowledged by L2 before the team can observe the arrival
    if (c.G > 1 && c.local) __builtin_amdgcn_s_waitcnt(WAIT_VMEM);
    __syncthreads();
    if (c.G > 1) {
        if (c.tid == 0) {
            if (!c.local) __threadfence();
            (void)__hip_atomic_fetch_add(c.bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long target = c.bar_next;
            unsigned spins = 0;
            while (__hip_atomic_load(c.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(2);
                // after one timeout anywhere in the team no barrier waits again,
                // so the launch drains quickly
                if (++spins > BAR_SPIN_LIMIT ||
                    __hip_atomic_load(c.bar_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    __hip_atomic_store(c.bar_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            c.bar_next = target + (unsigned long long)c.G;
            if (c.local) asm volatile("buffer_inv sc1" ::: "memory");  // L1 (no L2 write-back)
            else __threadfence();
        }
        __syncthreads();
        // profiling: team barriers per op, after the fixed slots
        if (c.prof && c.tid == 0)
            __hip_atomic_fetch_add(c.prof + c.pbase + PROF_EXTRA + c.cur_op, 1ull, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The sub-team of ranks [first, first + size) with its own barrier counter,
// reduction scratch and buffer parity, for work that runs beside the rest of
// the team.  Only rank 0 keeps the profile accumulators.
__device__ __forceinline__ Ctx sub_team(const Ctx& c, int first, int size, unsigned long long* bar, float* tred,
                                        int tpar) {
    Ctx s = c;
    s.G = size;
    s.rank = c.rank - first;
    s.gtid = s.rank * NT + c.tid;
    s.gsize = size * NT;
    s.bar = bar;
    if (size > 1 && c.tid == 0) s.bar_next = bar_first(bar, size);
    s.tred = tred;
    s.tpar = tpar;
    if (c.rank != 0) s.prof = nullptr;
    return s;
}

// The sub-team of ranks 1..G-1 (G >= 2), for work that runs beside rank 0.
__device__ __forceinline__ Ctx sub_team(const Ctx& c) { return sub_team(c, 1, c.G - 1, c.bar2, c.tred2, c.tpar2); }

// XCD of the calling workgroup (hardware register XCC_ID).
__device__ __forceinline__ unsigned xcd_id() { return __builtin_amdgcn_s_getreg((15 << 11) | (0 << 6) | 20); }

// At launch start of a team: every workgroup records its XCD; the team uses
// L2-local barriers only if all agree.  All threads of the team.
__device__ void team_place(Ctx& c, int* xtab) {
    c.local = 0;
    if (c.G == 1) return;
    if (c.tid == 0) __hip_atomic_store(&xtab[c.rank], (int)xcd_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    team_sync(c);
    int same = 1;
    int x0 = __hip_atomic_load(&xtab[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int r = 1; r < c.G; ++r) same &= __hip_atomic_load(&xtab[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == x0;
    team_sync(c);  // every workgroup has read the table before it is rewritten
    c.local = same;
    if (c.rank == 0 && c.tid == 0) xtab[-1] = same;  // reported to the host (barrier word 3)
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Block sum in a fixed order (lane tree, then waves in index order): the
// result is the same on every run.  All threads must call it.
__device__ float block_sum(Ctx& c, float v) {
    v = wave_sum(v);
    __syncthreads();
    if ((c.tid & 63) == 0) c.red[c.tid >> 6] = v;
    __syncthreads();
    float t = 0.f;
    for (int w = 0; w < NWAVE; ++w) t += c.red[w];
    __syncthreads();
    return t;
}
__device__ float block_max(Ctx& c, float v) {
    v = wave_max(v);
    __syncthreads();
    if ((c.tid & 63) == 0) c.red[c.tid >> 6] = v;
    __syncthreads();
    float t = c.red[0];
    for (int w = 1; w < NWAVE; ++w) t = fmaxf(t, c.red[w]);
    __syncthreads();
    return t;
}
__device__ int block_or(Ctx& c, int v) {
    int any = __syncthreads_or(v);
    return any;
}

// Team reductions: workgroup partials combined in rank order (fixed order, so
// the result is the same on every run).  All threads of the team.
__device__ float team_sum(Ctx& c, float v) {
    float t = block_sum(c, v);
    if (c.G == 1) return t;
    float* buf = c.tred + c.tpar * c.G;
    c.tpar ^= 1;
    if (c.tid == 0) buf[c.rank] = t;
    team_sync(c);
    float s = 0.f;
    for (int r = 0; r < c.G; ++r) s += buf[r];
    return s;
}
__device__ float team_max(Ctx& c, float v) {
    float t = block_max(c, v);
    if (c.G == 1) return t;
    float* buf = c.tred + c.tpar * c.G;
    c.tpar ^= 1;
    if (c.tid == 0) buf[c.rank] = t;
    team_sync(c);
    float s = buf[0];
    for (int r = 1; r < c.G; ++r) s = fmaxf(s, buf[r]);
    return s;
}
__device__ int team_or(Ctx& c, int v) {
    int t = __syncthreads_or(v);
    if (c.G == 1) return t;
    return team_max(c, t ? 1.f : 0.f) > 0.f ? 1 : 0;
}

// In-place exclusive scan of the ints a[0..n) (any memory); returns the
// total.  Each thread scans a contiguous chunk, chunk totals are scanned
// across the block.  All threads must call it.
template <class P>
__device__ int block_exscan(Ctx& c, P a, int n) {
    int per = (n + NT - 1) / NT;
    int b = c.tid * per, e = min(b + per, n);
    int s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    int lane = c.tid & 63, w = c.tid >> 6;
    int v = s;
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    __syncthreads();
    if (lane == 63) c.ired[w] = v;
    __syncthreads();
    int before = 0, tot = 0;
    for (int k = 0; k < NWAVE; ++k) {
        int x = c.ired[k];
        if (k < w) before += x;
        tot += x;
    }
    int run = before + v - s;
    for (int i = b; i < e; ++i) {
        int x = a[i];
        a[i] = run;
        run += x;
    }
    __syncthreads();
    return tot;
}

// Team-wide in-place exclusive scan of the ints a[0..n) in global memory;
// returns the total.  Each workgroup scans a contiguous chunk, chunk totals
// are added in rank order.  All threads of the team.
template <class P>
__device__ int team_exscan(Ctx& c, P a, int n) {
    if (c.G == 1) return block_exscan(c, a, n);
    int per = (n + c.G - 1) / c.G;
    int b = min(c.rank * per, n), e = min(b + per, n);
    int tot = block_exscan(c, a + b, e - b);
    float* buf = c.tred + c.tpar * c.G;
    c.tpar ^= 1;
    if (c.tid == 0) buf[c.rank] = __int_as_float(tot);
    team_sync(c);
    int before = 0, all = 0;
    for (int r = 0; r < c.G; ++r) {
        int x = __float_as_int(buf[r]);
        if (r < c.rank) before += x;
        all += x;
    }
    for (int i = b + c.tid; i < e; i += NT) a[i] += before;
    team_sync(c);
    return all;
}

// Sum over aligned groups of G lanes (G a power of two <= 64), fixed order;
// every lane of the group receives the group total.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---- splines (src/spline.h) -------------------------------------------------
// Uniform cubic B-spline with coefficient window p[xb-1 .. xb+2] at x>=1
// (first spline centred at -1, src/spline.h:98-127,242-256): value, derivative.
template <class P>
__device__ __forceinline__ void deboor(P p, float x, float& v, float& d) {
    int xb = (int)x;
    float y = x - (float)xb;
    P c = p + xb - 1;
    float c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    float omy = 1.f - y, y2 = y * y, y3 = y2 * y;
    float w0 = omy * omy * omy * (1.f / 6.f);
    float w1 = (3.f * y3 - 6.f * y2 + 4.f) * (1.f / 6.f);
    float w2 = (-3.f * y3 + 3.f * y2 + 3.f * y + 1.f) * (1.f / 6.f);
    float w3 = y3 * (1.f / 6.f);
    v = w0 * c0 + w1 * c1 + w2 * c2 + w3 * c3;
    float e0 = -0.5f * omy * omy;
    float e1 = 1.5f * y2 - 2.f * y;
    float e2 = -1.5f * y2 + y + 0.5f;
    float e3 = 0.5f * y2;
    d = e0 * c0 + e1 * c1 + e2 * c2 + e3 * c3;
}

// Clamped variant (src/spline.h:326-358): constant outside [1, n_knot-2).
template <class P>
__device__ __forceinline__ void clamped_deboor(P p, float x, int n_knot, float& v, float& d) {
    if (!(x >= 1.f)) {  // also catches NaN, which must never index the table
        v = p[0] * (1.f / 6.f) + p[1] * (2.f / 3.f) + p[2] * (1.f / 6.f);
        d = 0.f;
    } else if (x >= (float)(n_knot - 2)) {
        v = p[n_knot - 3] * (1.f / 6.f) + p[n_knot - 2] * (2.f / 3.f) + p[n_knot - 1] * (1.f / 6.f);
        d = 0.f;
    } else {
        deboor(p, x, v, d);
    }
}

// 1/(1+exp(x)) and its derivative exp(x)/(1+exp(x))^2 (src/vector_math.h:683-690).
__device__ __forceinline__ void sigmoid(float x, float& v, float& d) {
    float z = expf(-x);
    float w = 1.f / (1.f + z);
    v = w;
    d = z * w * w;
}

// Compact sigmoid 0.25(y+2)(y-1)^2 with y = x*sharpness, 1 below -1, 0 above 1
// (src/vector_math.h:701-720).
template <typename S>
__device__ __forceinline__ void compact_sigmoid(S x, S sharp, S& v, S& d) {
    S y = x * sharp;
    v = S(0.25) * (y + S(2)) * (y - S(1)) * (y - S(1));
    d = (sharp * S(0.75)) * (y * y - S(1));
    if (y < S(-1)) { v = S(1); d = S(0); }
    if (y > S(1)) { v = S(0); d = S(0); }
}

// Periodic bicubic patch c[16] (value, d/dfx, d/dfy), src/spline.h:63-82.
template <class P>
__device__ __forceinline__ void bicubic(P c, float fx, float fy, float& v, float& dx, float& dy) {
    float fx2 = fx * fx, fx3 = fx * fx2, fy2 = fy * fy;
    float vx0 = c[0] + fy * (c[1] + fy * (c[2] + fy * c[3]));
    float vx1 = c[4] + fy * (c[5] + fy * (c[6] + fy * c[7]));
    float vx2 = c[8] + fy * (c[9] + fy * (c[10] + fy * c[11]));
    float vx3 = c[12] + fy * (c[13] + fy * (c[14] + fy * c[15]));
    float vy1 = c[1] + fx * (c[5] + fx * (c[9] + fx * c[13]));
    float vy2 = c[2] + fx * (c[6] + fx * (c[10] + fx * c[14]));
    float vy3 = c[3] + fx * (c[7] + fx * (c[11] + fx * c[15]));
    dx = vx1 + 2.f * fx * vx2 + 3.f * fx2 * vx3;
    dy = vy1 + 2.f * fy * vy2 + 3.f * fy2 * vy3;
    v = vx0 + fx * vx1 + fx2 * vx2 + fx3 * vx3;
}

template <class P>
__device__ __forceinline__ void quat_to_rot(P q, float* U) {
    float a = q[0], b = q[1], c = q[2], d = q[3];
    U[0] = a * a + b * b - c * c - d * d; U[1] = 2.f * b * c - 2.f * a * d; U[2] = 2.f * b * d + 2.f * a * c;
    U[3] = 2.f * b * c + 2.f * a * d; U[4] = a * a - b * b + c * c - d * d; U[5] = 2.f * c * d - 2.f * a * b;
    U[6] = 2.f * b * d - 2.f * a * c; U[7] = 2.f * c * d + 2.f * a * b; U[8] = a * a - b * b - c * c + d * d;
}

// Fixed-order gather (see engine_dev.h): each target row adds the sum of its
// contributions, in plan order.  A thread loads each contribution's index
// once and all its components with it (per component the terms are summed in
// plan order, as one component at a time would).  Called by all threads.
template <int CW>
__device__ __forceinline__ void plan_gather_w(Ctx& c, gp<const int> off, gp<const int> idx, gp<float> sens,
                                              gp<const float> con, int nt, int ss, int col, int cs) {
    for (int t = c.gtid; t < nt; t += c.gsize) {
        const int b = off[t], e = off[t + 1];
        if (b == e) continue;
        float acc[CW];
#pragma unroll
        for (int k = 0; k < CW; ++k) acc[k] = 0.f;
#pragma unroll 2
        for (int j = b; j < e; ++j) {
            gp<const float> q = con + idx[j] * cs;
#pragma unroll
            for (int k = 0; k < CW; ++k) acc[k] += q[k];
        }
#pragma unroll
        for (int k = 0; k < CW; ++k) sens[t * ss + col + k] += acc[k];
    }
}

__device__ void plan_gather(Ctx& c, int plan) {
    if (plan < 0) return;
    gp<const int> h = gmem(c.CI + plan);
    int nt = h[0], cw = h[1], so = h[2], ss = h[3], col = h[4], cb = h[5], cs = h[6];
    gp<const int> off = h + PLAN_HDR;
    gp<const int> idx = off + nt + 1;
    gp<float> sens = gmem(c.A + so);
    gp<const float> con = gmem((const float*)c.A + cb);
    switch (cw) {
    case 1: plan_gather_w<1>(c, off, idx, sens, con, nt, ss, col, cs); break;
    case 2: plan_gather_w<2>(c, off, idx, sens, con, nt, ss, col, cs); break;
    case 3: plan_gather_w<3>(c, off, idx, sens, con, nt, ss, col, cs); break;
    case 6: plan_gather_w<6>(c, off, idx, sens, con, nt, ss, col, cs); break;
    default:
        for (int t = c.gtid; t < nt; t += c.gsize) {
            int b = off[t], e = off[t + 1];
            if (b == e) continue;
            for (int k = 0; k < cw; ++k) {
                float acc = 0.f;
                for (int j = b; j < e; ++j) acc += con[idx[j] * cs + k];
                sens[t * ss + col + k] += acc;
            }
        }
    }
    team_sync(c);
}

}  // namespace upside
