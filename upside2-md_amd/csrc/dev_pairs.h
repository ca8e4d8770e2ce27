// Pair graphs on the device: Verlet cache from LDS-tiled scans, exact edge
// refine, per-edge spline energies and the fixed-order sensitivity gather.
//
// Reference: PairlistComputation / InteractionGraph (src/interaction_graph.h).
// The edge set of a graph is {(i1,i2) acceptable : dist2 < cutoff^2} where
// dist2 is computed from the current packed coordinates exactly as the
// reference build evaluates it, fma(x,x, fma(z,z, y*y)) (the Float4 sum of
// src/interaction_graph.h:239 contracted by the compiler).  The cache (radius
// cutoff + buffer, rebuilt when any element drifts more than buffer/2,
// src/interaction_graph.h:51-169) only bounds which pairs are examined, so the
// edge set does not depend on how the cache was built.
//
// Rebuild: the other side is put into a hashed cell grid in LDS (or streamed
// in tiles when it does not fit) and every row scans it twice, once to count
// partners and, after a team scan of the counts, once to write them.  Row
// order is canonical (fixed cell-visit order, index order inside a cell).
// Transposed rows (the forward slots of each side-2 element) are placed by
// per-column counters and then ordered by slot, so every order is
// independent of scheduling.
#pragma once
#include "dev_common.h"

namespace upside {

__device__ __forceinline__ float refine_dist2(float dx, float dy, float dz) {
    return __fmaf_rn(dx, dx, __fmaf_rn(dz, dz, __fmul_rn(dy, dy)));
}

__device__ __forceinline__ bool pg_accept(const PairGraph& G, int a, int b) {
    if (G.excl < 0) return (a >> 4) != (b >> 4);
    int d = a - b;
    return d > G.excl || -d > G.excl;
}

__device__ __forceinline__ const float* pg_side2(const Ctx& c, const PairGraph& G) {
    return c.A + G.pos_pack + (G.sym ? 0 : G.n1 * 3);
}

// Copy the current coordinates of both sides into pos_pack and report whether
// the cache must be rebuilt (first use, or drift beyond half the buffer).
__device__ bool pg_pack_and_check(Ctx& c, const PairGraph& G) {
    gp<const int> loc1 = gmem(c.CI + G.loc1);
    gp<const int> loc2 = gmem(c.CI + G.loc2);
    gp<const float> P1 = gmem((const float*)c.A + G.out1);
    gp<const float> P2 = gmem((const float*)c.A + G.out2);
    gp<float> pack = gmem(c.A + G.pos_pack);
    gp<const float> cache = gmem((const float*)c.A + G.cache_pos);
    int n = G.sym ? G.n1 : G.n1 + G.n2;
    int valid = c.I[G.state];
    int moved = 0;
    for (int i = c.gtid; i < n; i += c.gsize) {
        gp<const float> src = (i < G.n1) ? P1 + loc1[i] * G.w1 : P2 + loc2[i - G.n1] * G.w2;
        float x = src[0], y = src[1], z = src[2];
        pack[i * 3 + 0] = x; pack[i * 3 + 1] = y; pack[i * 3 + 2] = z;
        float dx = x - cache[i * 3 + 0], dy = y - cache[i * 3 + 1], dz = z - cache[i * 3 + 2];
        if (!(dx * dx + dy * dy + dz * dz <= G.drift2)) moved = 1;  // NaN also forces a rebuild
    }
    int any = team_or(c, moved | (valid ? 0 : 1));
    return any != 0;
}

template <class P>
__device__ int lower_bound_i(P a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

// ---- cache rebuild -------------------------------------------------------------
// Rows are scanned by groups of SG lanes: the lanes of a group stride one
// query's candidates together and emit the hits in candidate order through a
// ballot, so a row costs ceil(candidates / SG) dependent steps instead of one
// step per candidate.
constexpr int SG = 8;

struct GroupLanes {
    int gl;                    // lane within the group
    unsigned long long gmask;  // the group's lanes in the wave
    unsigned long long below;  // wave lanes below this one
};
__device__ __forceinline__ GroupLanes group_lanes(const Ctx& c) {
    GroupLanes L;
    int lane = c.tid & 63;
    L.gl = lane & (SG - 1);
    L.gmask = ((1ull << SG) - 1ull) << (lane & ~(SG - 1));
    L.below = (1ull << lane) - 1ull;
    return L;
}

// One SG-wide step of a row: returns the group's hits; FILL writes each hit at
// cur + (hits of the lower lanes), so hits keep candidate order.  Every lane of
// the group calls it.
template <bool FILL>
__device__ __forceinline__ int group_emit(const GroupLanes& L, bool hit, int cur, int q, int j, gp<int> nbr,
                                         gp<int> row) {
    unsigned long long m = __ballot(hit) & L.gmask;
    if (FILL && hit) {
        int p = cur + __popcll(m & L.below);
        nbr[p] = j;
        row[p] = q;
    }
    return __popcll(m);
}

// Fallback scan for sides too large for the LDS grid: the other side streamed
// through LDS in tiles of TILE elements, partners in increasing index.
constexpr int TILE = LDS_FLOATS / 4;

// Elements staged in LDS are 4 floats: position and an integer tag stored as
// float bits (the four stores/loads combine into 16-byte LDS accesses).
__device__ __forceinline__ void el_put(lp<float> el, int k, float x, float y, float z, int tag) {
    el[4 * k] = x; el[4 * k + 1] = y; el[4 * k + 2] = z; el[4 * k + 3] = __int_as_float(tag);
}

// Stage elements [t0, t0+n) of side pts (ids in id) as (x, y, z, id).
__device__ void stage_tile(Ctx& c, gp<const float> pts, gp<const int> id, int t0, int n) {
    lp<float> tile = lmem(c.lds);
    for (int k = c.tid; k < n; k += NT) {  // every workgroup stages its own copy
        int j = t0 + k;
        el_put(tile, k, pts[j * 3], pts[j * 3 + 1], pts[j * 3 + 2], id[j]);
    }
    __syncthreads();
}

// Rows q in [0, nq) (positions qp, ids qid) against the elements ep[0..ne)
// (ids eid): a pair (q, j) is a hit when within the cache radius and
// acceptable (SYM: also j > q).  !FILL adds row q's hits to off[q].  FILL
// writes them at the cursor cur[q] (advanced).
template <bool SYM, bool FILL>
__device__ void tile_rows(Ctx& c, const PairGraph& G, gp<const float> qp, gp<const int> qid, int nq,
                          gp<const float> ep, gp<const int> eid, int ne, gp<int> off, gp<int> cur_arr, gp<int> nbr,
                          gp<int> row) {
    lp<const float> tile = (lp<const float>)lmem(c.lds);
    const float cc2 = G.cache_cut2;
    const GroupLanes L = group_lanes(c);
    const int grp = c.gtid / SG, ngrp = c.gsize / SG;
    for (int t0 = 0; t0 < ne; t0 += TILE) {
        int tn = min(TILE, ne - t0);
        stage_tile(c, ep, eid, t0, tn);
        for (int q = grp; q < nq; q += ngrp) {
            float x0 = qp[q * 3], x1 = qp[q * 3 + 1], x2 = qp[q * 3 + 2];
            int a = qid[q];
            int cur = FILL ? cur_arr[q] : 0;
            for (int k0 = 0; k0 < tn; k0 += SG) {
                int k = k0 + L.gl, j = t0 + k;
                bool hit = false;
                if (k < tn) {
                    float ex = tile[4 * k], ey = tile[4 * k + 1], ez = tile[4 * k + 2];
                    float d2 = refine_dist2(x0 - ex, x1 - ey, x2 - ez);
                    hit = d2 < cc2 && (!SYM || j > q) && pg_accept(G, a, __float_as_int(tile[4 * k + 3]));
                }
                cur += group_emit<FILL>(L, hit, cur, q, j, nbr, row);
            }
            if (L.gl == 0) {
                if (FILL) cur_arr[q] = cur;
                else off[q] += cur;
            }
        }
        __syncthreads();  // the tile is restaged next; cursors visible to the group
    }
    team_sync(c);
}

// ---- cache rebuild: cell grid in LDS ----------------------------------------
// The other side is counting-sorted into hashed cells of edge >= the cache
// radius, kept in LDS as (bucket offsets, elements (x, y, z, index),
// element ids), each bucket sorted by element index.  A row visits its 27
// neighbouring cells in a fixed order, so rows come out in a canonical order.
struct LGrid {
    float ox, oy, oz, inv;
    int mask;
    lp<int> ofs;      // [mask+2]
    lp<float> el;     // [4m] (x, y, z, element index)
    lp<int> eid;      // [m]  element ids, same order (no global loads in the scan)
};

__device__ __forceinline__ int cell_coord(float x, float o, float inv) {
    // clamp keeps far outliers (and NaN) inside int range; they only cost scan time
    float t = fminf(fmaxf((x - o) * inv, -1.0e6f), 1.0e6f);
    return (int)floorf(t);
}
__device__ __forceinline__ int cell_bucket(int mask, int cx, int cy, int cz) {
    uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return (int)(h & (uint32_t)mask);
}

__device__ __forceinline__ int lgrid_buckets(int m) {
    int nb = 64;
    while (nb < m && nb < 4096) nb *= 2;
    return nb;
}
__device__ __forceinline__ bool lgrid_fits(int m) {
    int nb = lgrid_buckets(m);
    return (nb + 4) * 4 + m * 20 <= LDS_FLOATS * 4;
}

// Build the grid of the m points pts (stride 3) in this workgroup's LDS
// (every workgroup of a team builds its own copy).  All threads of the workgroup.
__device__ LGrid lgrid_build(Ctx& c, const PairGraph& G, gp<const float> pts, gp<const int> ids, int m) {
    float mn0 = 3.0e38f, mn1 = 3.0e38f, mn2 = 3.0e38f;
    for (int i = c.tid; i < m; i += NT) {
        mn0 = fminf(mn0, pts[i * 3]);
        mn1 = fminf(mn1, pts[i * 3 + 1]);
        mn2 = fminf(mn2, pts[i * 3 + 2]);
    }
    LGrid g;
    g.ox = -block_max(c, -mn0);
    g.oy = -block_max(c, -mn1);
    g.oz = -block_max(c, -mn2);
    // cells slightly larger than the cache radius: rounding can never put two
    // points closer than the radius more than one cell apart
    g.inv = 1.f / (G.cache_cut * 1.0001f);
    int nb = lgrid_buckets(m);
    g.mask = nb - 1;
    lp<float> L = lmem(c.lds);
    g.ofs = (lp<int>)L;
    g.el = L + ((nb + 4) & ~3);
    g.eid = (lp<int>)(g.el + 4 * m);
    for (int b = c.tid; b <= nb; b += NT) g.ofs[b] = 0;
    __syncthreads();
    for (int i = c.tid; i < m; i += NT) {
        int b = cell_bucket(g.mask, cell_coord(pts[i * 3], g.ox, g.inv), cell_coord(pts[i * 3 + 1], g.oy, g.inv),
                            cell_coord(pts[i * 3 + 2], g.oz, g.inv));
        atomicAdd((int*)(g.ofs + b + 1), 1);
    }
    __syncthreads();
    block_exscan(c, (int*)(g.ofs + 1), nb);  // ofs[b+1] = start of bucket b
    for (int i = c.tid; i < m; i += NT) {
        float x = pts[i * 3], y = pts[i * 3 + 1], z = pts[i * 3 + 2];
        int b = cell_bucket(g.mask, cell_coord(x, g.ox, g.inv), cell_coord(y, g.oy, g.inv), cell_coord(z, g.oz, g.inv));
        int p = atomicAdd((int*)(g.ofs + b + 1), 1);  // ends at the start of bucket b+1
        el_put(g.el, p, x, y, z, i);
        g.eid[p] = ids[i];
    }
    __syncthreads();
    // canonical order inside every bucket: by element index
    for (int b = c.tid; b < nb; b += NT) {
        int lo = g.ofs[b], hi = g.ofs[b + 1];
        for (int k = lo + 1; k < hi; ++k) {
            float vx = g.el[4 * k], vy = g.el[4 * k + 1], vz = g.el[4 * k + 2];
            int key = __float_as_int(g.el[4 * k + 3]);
            int vid = g.eid[k];
            int t = k - 1;
            while (t >= lo && __float_as_int(g.el[4 * t + 3]) > key) {
#pragma unroll
                for (int q = 0; q < 4; ++q) g.el[4 * (t + 1) + q] = g.el[4 * t + q];
                g.eid[t + 1] = g.eid[t];
                --t;
            }
            el_put(g.el, t + 1, vx, vy, vz, key);
            g.eid[t + 1] = vid;
        }
    }
    __syncthreads();
    return g;
}

// Rows against the grid; same hit rule as tile_rows.  !FILL: off[q] = hits of
// row q.  FILL: row q's hits are written from off[q] on.
template <bool SYM, bool FILL>
__device__ void grid_rows(Ctx& c, const PairGraph& G, const LGrid& g, gp<const float> qp, gp<const int> qid, int nq,
                          gp<int> off, gp<int> nbr, gp<int> row) {
    const float cc2 = G.cache_cut2;
    const GroupLanes L = group_lanes(c);
    const int grp = c.gtid / SG, ngrp = c.gsize / SG;
    for (int q = grp; q < nq; q += ngrp) {
        float x0 = qp[q * 3], x1 = qp[q * 3 + 1], x2 = qp[q * 3 + 2];
        int a = qid[q];
        int cur = FILL ? off[q] : 0;
        int cx = cell_coord(x0, g.ox, g.inv), cy = cell_coord(x1, g.oy, g.inv), cz = cell_coord(x2, g.oz, g.inv);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int ax = cx + dx, ay = cy + dy, az = cz + dz;
                    int b = cell_bucket(g.mask, ax, ay, az);
                    int t1 = g.ofs[b + 1];
                    for (int t = g.ofs[b]; t < t1; t += SG) {
                        int k = t + L.gl, j = 0;
                        bool hit = false;
                        if (k < t1) {
                            float ex = g.el[4 * k], ey = g.el[4 * k + 1], ez = g.el[4 * k + 2];
                            j = __float_as_int(g.el[4 * k + 3]);
                            float d2 = refine_dist2(x0 - ex, x1 - ey, x2 - ez);
                            // the cell test drops elements of other cells sharing the bucket
                            hit = d2 < cc2 && (!SYM || j > q) && cell_coord(ex, g.ox, g.inv) == ax &&
                                  cell_coord(ey, g.oy, g.inv) == ay && cell_coord(ez, g.oz, g.inv) == az &&
                                  pg_accept(G, a, g.eid[k]);
                        }
                        cur += group_emit<FILL>(L, hit, cur, q, j, nbr, row);
                    }
                }
        if (!FILL && L.gl == 0) off[q] = cur;
    }
    team_sync(c);
}

// Order every transposed row by forward slot: entry tmp[t] of column j (the
// slots whose partner is j, in arrival order) goes to tslot[toff[j] + r],
// r = number of entries of the column with a smaller slot.  Columns are split
// among the team's workgroups; each workgroup stages runs of whole columns in
// LDS.  All threads of the team.
__device__ void pg_order_columns(Ctx& c, const PairGraph& G, int n2) {
    gp<const int> toff = gmem((const int*)c.I + G.toff);
    gp<const int> tmp = gmem((const int*)c.I + G.tmp);
    gp<const int> nbr = gmem((const int*)c.I + G.nbr);
    gp<int> tslot = gmem(c.I + G.tslot);
    lp<int> sl = (lp<int>)lmem(c.lds);
    const int CAP = LDS_FLOATS;
    int per = (n2 + c.G - 1) / c.G;
    int jb = min(c.rank * per, n2);
    const int jend = min(jb + per, n2);
    while (jb < jend) {
        const int base = toff[jb];
        // the longest run of whole columns [jb, je) that fits: largest je in
        // [jb+1, jend] with toff[je] - base <= CAP (je = jb+1 if none)
        int lo = jb + 1, hi = jend;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (toff[mid] - base <= CAP) lo = mid; else hi = mid - 1;
        }
        const int je = lo;
        const int n = toff[je] - base;
        if (n <= CAP) {
            for (int k = c.tid; k < n; k += NT) sl[k] = tmp[base + k];
            __syncthreads();
            for (int k = c.tid; k < n; k += NT) {
                int s = sl[k];
                int j = nbr[s];
                int u0 = toff[j] - base, u1 = toff[j + 1] - base;
                int r = 0;
                for (int u = u0; u < u1; ++u) r += sl[u] < s;
                tslot[base + u0 + r] = s;
            }
            __syncthreads();
        } else {  // one column longer than the LDS area: rank from memory
            for (int k = c.tid; k < n; k += NT) {
                int s = tmp[base + k];
                int r = 0;
                for (int u = 0; u < n; ++u) r += tmp[base + u] < s;
                tslot[base + r] = s;
            }
        }
        jb = je;
    }
    team_sync(c);
}

// Rebuild rows and transposed rows from pos_pack.  Sets state[1] on overflow
// (the host grows the capacity and repeats the run).
//   forward rows:    side-1 element i, partners j (symmetric: j > i) in the
//                    canonical scan order;
//   transposed rows: side-2 element j, the forward slots with partner j in
//                    increasing slot order.
__device__ void pg_rebuild(Ctx& c, const PairGraph& G) {
    gp<const int> id1 = gmem(c.CI + G.id1);
    const bool sym = G.sym != 0;
    gp<const int> id2 = sym ? id1 : gmem(c.CI + G.id2);
    gp<const float> pack = gmem((const float*)c.A + G.pos_pack);
    gp<float> cache = gmem(c.A + G.cache_pos);
    gp<const float> p1 = pack;
    gp<const float> p2 = gmem(pg_side2(c, G));
    const int n1 = G.n1, n2 = sym ? G.n1 : G.n2;
    gp<int> off = gmem(c.I + G.off);
    gp<int> nbr = gmem(c.I + G.nbr);
    gp<int> row = gmem(c.I + G.row);
    gp<int> toff = gmem(c.I + G.toff);
    gp<int> tmp = gmem(c.I + G.tmp);
    gp<int> tcur = gmem(c.I + G.tcur);
    int npack = sym ? n1 : n1 + n2;
    for (int i = c.gtid; i < npack * 3; i += c.gsize) cache[i] = pack[i];
    for (int i = c.gtid; i <= n1; i += c.gsize) off[i] = 0;
    for (int j = c.gtid; j <= n2; j += c.gsize) toff[j] = 0;
    team_sync(c);
    unsigned long long t0 = ptick(c);
    // Small graphs scan every row against the whole other side staged in LDS
    // (n1 * n2 distance tests spread over the team cost less than the grid
    // build's serial phases; measured: 4-5 % of a step at 150-500 residues);
    // larger ones use the cell grid, which stays in LDS from the count scan
    // to the fill scan.  The choice depends on the sizes only, so every team
    // size builds the same rows.
    const bool use_grid = lgrid_fits(n2) && (long long)n1 * n2 > (1ll << 21);
    LGrid g{};
    if (use_grid) {
        g = lgrid_build(c, G, p2, id2, n2);
        pacc(c, PROF_RB_GRID, t0);
    }
    unsigned long long t1 = ptick(c);
    if (use_grid) {
        if (sym) grid_rows<true, false>(c, G, g, p1, id1, n1, off, off, off);
        else grid_rows<false, false>(c, G, g, p1, id1, n1, off, off, off);
    } else {
        if (sym) tile_rows<true, false>(c, G, p1, id1, n1, p2, id2, n2, off, off, off, off);
        else tile_rows<false, false>(c, G, p1, id1, n1, p2, id2, n2, off, off, off, off);
    }
    pacc(c, PROF_RB_SCAN, t1);
    const int E = team_exscan(c, (int*)off, n1 + 1);
    if (E > G.ecap) {
        if (c.gtid == 0) { c.I[G.state + 1] = 1; c.I[G.state + 0] = 0; }
        team_sync(c);
        return;
    }
    pacc(c, PROF_RB_COUNT, t0);
    t0 = ptick(c);
    t1 = ptick(c);
    if (use_grid) {
        if (sym) grid_rows<true, true>(c, G, g, p1, id1, n1, off, nbr, row);
        else grid_rows<false, true>(c, G, g, p1, id1, n1, off, nbr, row);
    } else {
        for (int i = c.gtid; i < n1; i += c.gsize) tcur[i] = off[i];
        team_sync(c);
        if (sym) tile_rows<true, true>(c, G, p1, id1, n1, p2, id2, n2, off, tcur, nbr, row);
        else tile_rows<false, true>(c, G, p1, id1, n1, p2, id2, n2, off, tcur, nbr, row);
    }
    pacc(c, PROF_RB_SCAN, t1);
    pacc(c, PROF_RB_FILL, t0);
    t0 = ptick(c);
    // transposed rows: column counts, arrival-order placement, slot order
    for (int s = c.gtid; s < E; s += c.gsize) atomicAdd((int*)(toff + nbr[s]), 1);
    team_sync(c);
    team_exscan(c, (int*)toff, n2 + 1);
    for (int j = c.gtid; j < n2; j += c.gsize) tcur[j] = toff[j];
    team_sync(c);
    for (int s = c.gtid; s < E; s += c.gsize) tmp[atomicAdd((int*)(tcur + nbr[s]), 1)] = s;
    team_sync(c);
    pg_order_columns(c, G, n2);
    pacc(c, PROF_RB_MAP, t0);
    if (c.gtid == 0) {
        c.I[G.state + 0] = 1;
        c.I[G.state + 2] += 1;
        c.I[G.state + 3] = E;
    }
    team_sync(c);
}

// ---- edge functions -----------------------------------------------------
// quadspline (src/bead_interaction.h:31-85): radial clamped splines times
// angular splines of the two bead directions.
template <int NKA, int NK, class P, class X>
__device__ __forceinline__ float quadspline(P p, X x1, X x2, float inv_dtheta, float inv_dx, float* d1, float* d2) {
    float dx = x2[0] - x1[0], dy = x2[1] - x1[1], dz = x2[2] - x1[2];
    float r1x = x1[3], r1y = x1[4], r1z = x1[5];
    float r2x = x2[3], r2y = x2[4], r2z = x2[5];
    float dist2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, dx * dx));
    float inv = rsqrtf(dist2);
    float dist_coord = dist2 * (inv * inv_dx);
    float ux = inv * dx, uy = inv * dy, uz = inv * dz;
    float c1 = r1x * ux + r1y * uy + r1z * uz;
    float c2 = -(r2x * ux + r2y * uy + r2z * uz);
    // the window p[xb-1..xb+2] must stay inside the parameter row (NaN -> 1)
    float xa1 = fminf(fmaxf((c1 + 1.f) * inv_dtheta + 1.f, 1.f), (float)(NKA - 2));
    float xa2 = fminf(fmaxf((c2 + 1.f) * inv_dtheta + 1.f, 1.f), (float)(NKA - 2));
    float a1, da1, a2, da2, w, dw, nn, dn;
    deboor(p, xa1, a1, da1);
    deboor(p + NKA, xa2, a2, da2);
    clamped_deboor(p + 2 * NKA, dist_coord, NK, w, dw);
    clamped_deboor(p + 2 * NKA + NK, dist_coord, NK, nn, dn);
    float aw = a1 * a2;
    float rad = inv_dx * (dw + aw * dn);
    float ad1 = inv_dtheta * da1 * a2 * nn;
    float ad2 = inv_dtheta * a1 * da2 * nn;
    float rx = ad1 * r1x - ad2 * r2x, ry = ad1 * r1y - ad2 * r2y, rz = ad1 * r1z - ad2 * r2z;
    float ur = ux * rx + uy * ry + uz * rz;
    float ddx = rad * ux + inv * (rx - ur * ux);
    float ddy = rad * uy + inv * (ry - ur * uy);
    float ddz = rad * uz + inv * (rz - ur * uz);
    d1[0] = -ddx; d1[1] = -ddy; d1[2] = -ddz;
    d1[3] = ad1 * ux; d1[4] = ad1 * uy; d1[5] = ad1 * uz;
    d2[0] = ddx; d2[1] = ddy; d2[2] = ddz;
    d2[3] = -ad2 * ux; d2[4] = -ad2 * uy; d2[5] = -ad2 * uz;
    return w + aw * nn;
}

// ProteinHBondInteraction (src/hbond.cpp:158-246).
template <class P, class X>
__device__ __forceinline__ float protein_hbond_edge(P p, X x1, X x2, float* d1, float* d2) {
    float HOx = x1[0] - x2[0], HOy = x1[1] - x2[1], HOz = x1[2] - x2[2];
    float m2 = __fmaf_rn(HOz, HOz, __fmaf_rn(HOy, HOy, HOx * HOx)) + 1e-6f;
    float inv = rsqrtf(m2);
    float mag = m2 * inv;
    float rx = HOx * inv, ry = HOy * inv, rz = HOz * inv;
    float dHOC = rx * x2[3] + ry * x2[4] + rz * x2[5];
    float dOHN = -(rx * x1[3] + ry * x1[4] + rz * x1[5]);
    float hb = 0.f, dH[3] = {0, 0, 0}, drHN[3] = {0, 0, 0}, drOC[3] = {0, 0, 0};
    if (dHOC > 0.f && dOHN > 0.f) {
        float os, osd, is, isd, a1, da1, a2, da2;
        sigmoid((p[2] - mag) * p[3], os, osd);
        sigmoid((mag - p[0]) * p[1], is, isd);
        float rv = os * is;
        float rd = -p[3] * osd * is + p[1] * isd * os;
        sigmoid((dHOC - p[4]) * p[5], a1, da1); da1 *= p[5];
        sigmoid((dOHN - p[4]) * p[5], a2, da2); da2 *= p[5];
        hb = rv * a1 * a2;
        float c0 = rd * a1 * a2, c1 = rv * da1 * a2, c2 = -rv * a1 * da2;
        float r[3] = {rx, ry, rz};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            drOC[k] = c1 * r[k];
            drHN[k] = c2 * r[k];
            dH[k] = c0 * r[k] + (c1 * inv) * (x2[3 + k] - dHOC * r[k]) + (c2 * inv) * (x1[3 + k] + dOHN * r[k]);
        }
    }
    float hb_log = (1.f <= hb) ? 100.f : -logf(1.f - hb);
    float pref = fminf(1.f / (1.f - hb), 1e5f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d1[k] = dH[k] * pref; d1[3 + k] = drHN[k] * pref;
        d2[k] = -dH[k] * pref; d2[3 + k] = drOC[k] * pref;
    }
    return hb_log;
}

// Environment coverage (src/environment.cpp:114-228); weighted: second element
// carries a probability in component 3.
template <bool WEIGHTED, class P, class X>
__device__ __forceinline__ float env_edge(P p, X x1, X x2, float* d1, float* d2) {
    float dx = x2[0] - x1[0], dy = x2[1] - x1[1], dz = x2[2] - x1[2];
    float dist2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, dx * dx));
    float inv = rsqrtf(dist2);
    float dist = dist2 * inv;
    float ux = inv * dx, uy = inv * dy, uz = inv * dz;
    float dp = ux * x1[3] + uy * x1[4] + uz * x1[5];
    float rs, drs, as, das;
    compact_sigmoid(dist - p[0], p[1], rs, drs);
    compact_sigmoid(p[2] - dp, p[3], as, das);
    float prob = WEIGHTED ? x2[3] : 1.f;
    float u[3] = {ux, uy, uz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float ddk = prob * (drs * as * u[k] - rs * das * inv * (x1[3 + k] - dp * u[k]));
        d1[k] = -ddk;
        d1[3 + k] = -prob * rs * das * u[k];
        d2[k] = ddk;
    }
    float score = rs * as;
    if (WEIGHTED) {
        d2[3] = score;
        return prob * score;
    }
    d2[3] = d2[4] = d2[5] = 0.f;
    return score;
}

// BackbonePairs residue pair (src/backbone_steric.cpp:94-140): compact-sigmoid
// repulsion between the placed backbone atoms (at most 4) of two residues;
// derivatives are (force, torque) on each residue frame.
__device__ float backbone_edge(const Ctx& c, const OpRec& op, int r1, int r2, float* d1, float* d2) {
    gp<const float> aff = gmem((const float*)c.A + op.in_out[0]);
    int w = op.in_w[0];
    gp<const float> ref = gmem(c.CF + op.ia[1]);
    gp<const int> nat = gmem(c.CI + op.ia[2]);
    gp<const int> res = gmem(c.CI + op.ia[0]);
    int R1 = res[r1], R2 = res[r2];
    float a1[7], a2[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) { a1[k] = aff[R1 * w + k]; a2[k] = aff[R2 * w + k]; }
    float U1[9], U2[9];
    quat_to_rot(a1 + 3, U1);
    quat_to_rot(a2 + 3, U2);
    float X1[4][3], X2[4][3];
    int n1 = nat[r1], n2 = nat[r2];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gp<const float> q = ref + (r1 * 4 + a) * 3;
        gp<const float> s = ref + (r2 * 4 + a) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            X1[a][k] = U1[3 * k] * q[0] + U1[3 * k + 1] * q[1] + U1[3 * k + 2] * q[2] + a1[k];
            X2[a][k] = U2[3 * k] * s[0] + U2[3 * k + 1] * s[1] + U2[3 * k + 2] * s[2] + a2[k];
        }
    }
    const float cut2 = 3.f * 3.f + 0.1f * 3.f;
    const float sharp = 1.f / (3.0f * 0.10f);
    float pot = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) { d1[k] = 0.f; d2[k] = 0.f; }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i >= n1 || j >= n2) continue;
            float rx = X1[i][0] - X2[j][0], ry = X1[i][1] - X2[j][1], rz = X1[i][2] - X2[j][2];
            float rm2 = rx * rx + ry * ry + rz * rz;
            if (rm2 > cut2) continue;
            float v, dv;
            compact_sigmoid(rm2 - 9.f, sharp, v, dv);
            pot += 4.f * v;
            float g = 2.f * 4.f * dv;
            float gx = g * rx, gy = g * ry, gz = g * rz;
            float p1x = X1[i][0] - a1[0], p1y = X1[i][1] - a1[1], p1z = X1[i][2] - a1[2];
            float p2x = X2[j][0] - a2[0], p2y = X2[j][1] - a2[1], p2z = X2[j][2] - a2[2];
            d1[0] += gx; d1[1] += gy; d1[2] += gz;
            d1[3] += p1y * gz - p1z * gy; d1[4] += p1z * gx - p1x * gz; d1[5] += p1x * gy - p1y * gx;
            d2[0] -= gx; d2[1] -= gy; d2[2] -= gz;
            d2[3] -= p2y * gz - p2z * gy; d2[4] -= p2z * gx - p2x * gz; d2[5] -= p2x * gy - p2y * gx;
        }
    }
    return pot;
}

// Refine every cached slot against the cutoff and evaluate the active edges;
// param is the coefficient table (LDS or memory).
template <int KIND, int DD1, int DD2, class P>
__device__ void pg_edges_impl(Ctx& c, const PairGraph& G, const OpRec& op, P param) {
    gp<const int> nbr = gmem((const int*)c.I + G.nbr);
    gp<const int> row = gmem((const int*)c.I + G.row);
    gp<int> act = gmem(c.I + G.act);
    gp<float> val = gmem(c.A + G.val);
    gp<float> D1 = gmem(c.A + G.d1);
    gp<float> D2 = gmem(c.A + G.d2);
    gp<const float> pack = gmem((const float*)c.A + G.pos_pack);
    gp<const float> pk2 = gmem(pg_side2(c, G));
    gp<const int> loc1 = gmem(c.CI + G.loc1);
    gp<const int> loc2 = gmem(c.CI + G.loc2);
    gp<const int> t1 = gmem(c.CI + G.type1);
    gp<const int> t2 = gmem(c.CI + G.type2);
    gp<const float> P1 = gmem((const float*)c.A + G.out1);
    gp<const float> P2 = gmem((const float*)c.A + G.out2);
    const int E = c.I[G.state + 3];
    const int cap = G.ecap;
    for (int s = c.gtid; s < E; s += c.gsize) {
        int i1 = row[s], j = nbr[s];
        gp<const float> xp = pack + i1 * 3;
        gp<const float> yp = pk2 + j * 3;
        float d2 = refine_dist2(xp[0] - yp[0], xp[1] - yp[1], xp[2] - yp[2]);
        int a = d2 < G.cutoff2;
        act[s] = a;
        if (!a) { val[s] = 0.f; continue; }
        gp<const float> x1 = P1 + loc1[i1] * G.w1;
        gp<const float> x2 = P2 + loc2[j] * G.w2;
        P p = param + (t1[i1] * G.n_type2 + t2[j]) * G.n_param;
        float e1[7], e2[7];
        float v = 0.f;
        if (KIND == EDGE_ROTAMER) {
            v = quadspline<15, 12>(p, x1, x2, 6.f, 1.f / 0.7f, e1, e2);
        } else if (KIND == EDGE_HBOND_COVERAGE) {
            float cov = quadspline<15, 10>(p, x1, x2, 6.f, 1.f / 0.625f, e1, e2);
            float om = 1.f - x1[6];
            float pref = om * om;
#pragma unroll
            for (int q = 0; q < 6; ++q) { e1[q] *= pref; e2[q] *= pref; }
            e1[6] = -cov * om * 2.f;
            v = pref * cov;
        } else if (KIND == EDGE_PROTEIN_HBOND) {
            v = protein_hbond_edge(p, x1, x2, e1, e2);
        } else if (KIND == EDGE_ENV_WEIGHTED) {
            v = env_edge<true>(p, x1, x2, e1, e2);
        } else if (KIND == EDGE_ENV_HBOND) {
            v = env_edge<false>(p, x1, x2, e1, e2);
        } else {
            v = backbone_edge(c, op, i1, j, e1, e2);
        }
        val[s] = v;
#pragma unroll
        for (int q = 0; q < DD1; ++q) D1[q * cap + s] = e1[q];
#pragma unroll
        for (int q = 0; q < DD2; ++q) D2[q * cap + s] = e2[q];
    }
    team_sync(c);
}

// Stage the graph's coefficient table in this workgroup's LDS when it has one
// that fits, then evaluate.
template <int KIND, int DD1, int DD2>
__device__ void pg_edges_t(Ctx& c, const PairGraph& G, const OpRec& op) {
    gp<const float> src = gmem(c.CF + G.param);
    if (G.param_len > 0) {
        lp<float> L = lmem(c.lds);
        for (int k = c.tid; k < G.param_len; k += NT) L[k] = src[k];
        __syncthreads();
        pg_edges_impl<KIND, DD1, DD2>(c, G, op, (lp<const float>)L);
    } else {
        pg_edges_impl<KIND, DD1, DD2>(c, G, op, src);
    }
}

__device__ void pg_edges(Ctx& c, const PairGraph& G, const OpRec& op) {
    switch (G.kind) {
    case EDGE_ROTAMER: pg_edges_t<EDGE_ROTAMER, 6, 6>(c, G, op); break;
    case EDGE_HBOND_COVERAGE: pg_edges_t<EDGE_HBOND_COVERAGE, 7, 6>(c, G, op); break;
    case EDGE_PROTEIN_HBOND: pg_edges_t<EDGE_PROTEIN_HBOND, 6, 6>(c, G, op); break;
    case EDGE_ENV_WEIGHTED: pg_edges_t<EDGE_ENV_WEIGHTED, 6, 4>(c, G, op); break;
    case EDGE_ENV_HBOND: pg_edges_t<EDGE_ENV_HBOND, 6, 6>(c, G, op); break;
    default: pg_edges_t<EDGE_BACKBONE, 6, 6>(c, G, op); break;
    }
}

// Gather esens*deriv into the producers in a fixed order: groups of 16 lanes
// per element, each lane striding the element's slots, then a fixed-order
// group reduction.
constexpr int GW = 16;
static_assert(GW <= 32 && (GW & (GW - 1)) == 0, "a lane group is a power of two within half a wave");

template <int D1N, int D2N>
__device__ void pg_gather_t(Ctx& c, const PairGraph& G) {
    gp<const int> off = gmem((const int*)c.I + G.off);
    gp<const int> toff = gmem((const int*)c.I + G.toff);
    gp<const int> tslot = gmem((const int*)c.I + G.tslot);
    gp<const int> act = gmem((const int*)c.I + G.act);
    gp<const float> es = gmem((const float*)c.A + G.esens);
    gp<const float> D1 = gmem((const float*)c.A + G.d1);
    gp<const float> D2 = gmem((const float*)c.A + G.d2);
    gp<const int> loc1 = gmem(c.CI + G.loc1);
    gp<const int> loc2 = gmem(c.CI + G.loc2);
    gp<float> S1 = gmem(c.A + G.sens1);
    gp<float> S2 = gmem(c.A + G.sens2);
    const int cap = G.ecap;
    int grp = c.gtid / GW, gl = c.gtid % GW;
    const int NG = c.gsize / GW;
    // uniform trip counts: every lane runs every group reduction
    for (int base = 0; base < G.n1; base += NG) {
        int i = base + grp;
        bool ok = i < G.n1;
        float acc[D1N];
#pragma unroll
        for (int q = 0; q < D1N; ++q) acc[q] = 0.f;
        int b = ok ? off[i] : 0, e = ok ? off[i + 1] : 0;
        // branch-free: every load of an iteration is issued at once
#pragma unroll 2
        for (int s = b + gl; s < e; s += GW) {
            bool a = act[s] != 0;
            float w = es[s];
#pragma unroll
            for (int q = 0; q < D1N; ++q) {
                float d = D1[q * cap + s];
                acc[q] += a ? w * d : 0.f;
            }
        }
        if (G.sym && ok) {
            int tb = toff[i], te = toff[i + 1];
#pragma unroll 2
            for (int t = tb + gl; t < te; t += GW) {
                int s = tslot[t];
                bool a = act[s] != 0;
                float w = es[s];
#pragma unroll
                for (int q = 0; q < D1N; ++q) {
                    float d = D2[q * cap + s];
                    acc[q] += a ? w * d : 0.f;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < D1N; ++q) acc[q] = group_sum<GW>(acc[q]);
        if (ok && gl == 0) {
            gp<float> dst = S1 + loc1[i] * G.w1;
#pragma unroll
            for (int q = 0; q < D1N; ++q) dst[q] += acc[q];
        }
    }
    // both sides in one phase unless they accumulate into shared producer rows
    if (G.sym || G.gather_split) team_sync(c);
    if (G.sym) return;
    for (int base = 0; base < G.n2; base += NG) {
        int j = base + grp;
        bool ok = j < G.n2;
        float acc[D2N];
#pragma unroll
        for (int q = 0; q < D2N; ++q) acc[q] = 0.f;
        int tb = ok ? toff[j] : 0, te = ok ? toff[j + 1] : 0;
#pragma unroll 2
        for (int t = tb + gl; t < te; t += GW) {
            int s = tslot[t];
            bool a = act[s] != 0;
            float w = es[s];
#pragma unroll
            for (int q = 0; q < D2N; ++q) {
                float d = D2[q * cap + s];
                acc[q] += a ? w * d : 0.f;
            }
        }
#pragma unroll
        for (int q = 0; q < D2N; ++q) acc[q] = group_sum<GW>(acc[q]);
        if (ok && gl == 0) {
            gp<float> dst = S2 + loc2[j] * G.w2;
#pragma unroll
            for (int q = 0; q < D2N; ++q) dst[q] += acc[q];
        }
    }
    team_sync(c);
}

__device__ void pg_gather(Ctx& c, const PairGraph& G) {
    unsigned long long t0 = ptick(c);
    switch (G.kind) {
    case EDGE_HBOND_COVERAGE: pg_gather_t<7, 6>(c, G); break;
    case EDGE_ENV_WEIGHTED: pg_gather_t<6, 4>(c, G); break;
    default: pg_gather_t<6, 6>(c, G); break;  // symmetric graphs: both sides 6 wide
    }
    pacc(c, PROF_PG_GATHER, t0);
}

// Sum of the active edge values of each row (side 1) or transposed row
// (side 2), in slot order: out[k*stride] = sum.  Groups of 16 lanes.
template <class O>
__device__ void pg_row_sums(Ctx& c, const PairGraph& G, bool side2, O out, int stride) {
    gp<const int> off = gmem((const int*)c.I + (side2 ? G.toff : G.off));
    gp<const int> tslot = gmem((const int*)c.I + G.tslot);
    gp<const float> val = gmem((const float*)c.A + G.val);
    int n = side2 ? G.n2 : G.n1;
    int grp = c.gtid / GW, gl = c.gtid % GW;
    for (int base = 0; base < n; base += c.gsize / GW) {
        int r = base + grp;
        bool ok = r < n;
        float acc = 0.f;
        int k0 = ok ? off[r] : 0, k1 = ok ? off[r + 1] : 0;
        for (int k = k0 + gl; k < k1; k += GW) {
            int s = side2 ? tslot[k] : k;
            float v = val[s];  // zero on inactive slots
            acc += v;
        }
        acc = group_sum<GW>(acc);
        if (ok && gl == 0) out[r * stride] = acc;
    }
}

// Prepare a graph for an evaluation: pack, maybe rebuild, refine and evaluate.
// Returns false when the slot capacity overflowed (the evaluation is void).
__device__ bool pg_refresh(Ctx& c, const PairGraph& G, const OpRec& op) {
    unsigned long long t0 = ptick(c);
    if (pg_pack_and_check(c, G)) pg_rebuild(c, G);
    pacc(c, PROF_PG_REBUILD, t0);
    if (!c.I[G.state + 0]) return false;
    t0 = ptick(c);
    pg_edges(c, G, op);
    pacc(c, PROF_PG_EDGES, t0);
    return true;
}

}  // namespace upside
