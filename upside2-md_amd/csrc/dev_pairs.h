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
// Rebuild (all in one workgroup): the other side is put into a hashed cell
// grid in LDS (or streamed in tiles when it does not fit) and every row scans
// it twice, once to count partners and, after a block scan of the counts,
// once to write them.  Row order is canonical (fixed cell-visit order, index
// order inside a cell); transposed rows are scanned the same way from the
// other side and mapped to forward slots through a hash table.
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
    const int* loc1 = c.CI + G.loc1;
    const int* loc2 = c.CI + G.loc2;
    const float* P1 = c.A + G.out1;
    const float* P2 = c.A + G.out2;
    float* pack = c.A + G.pos_pack;
    const float* cache = c.A + G.cache_pos;
    int n = G.sym ? G.n1 : G.n1 + G.n2;
    int valid = c.I[G.state];
    int moved = 0;
    for (int i = c.gtid; i < n; i += c.gsize) {
        const float* src = (i < G.n1) ? P1 + loc1[i] * G.w1 : P2 + loc2[i - G.n1] * G.w2;
        float x = src[0], y = src[1], z = src[2];
        pack[i * 3 + 0] = x; pack[i * 3 + 1] = y; pack[i * 3 + 2] = z;
        float dx = x - cache[i * 3 + 0], dy = y - cache[i * 3 + 1], dz = z - cache[i * 3 + 2];
        if (!(dx * dx + dy * dy + dz * dz <= G.drift2)) moved = 1;  // NaN also forces a rebuild
    }
    int any = team_or(c, moved | (valid ? 0 : 1));
    return any != 0;
}

__device__ int lower_bound_i(const int* a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

// ---- cache rebuild: LDS-tiled scans ------------------------------------------
// Fallback scan: the other side streamed through LDS in tiles; every wave
// reads the same tile entries (LDS broadcast), each lane tests its own query.
// Partners come out in increasing index.
constexpr int TILE = LDS_FLOATS / 4;

// Stage elements [t0, t0+n) of side pts (ids in id) as float4 (x, y, z, id).
__device__ void stage_tile(Ctx& c, const float* pts, const int* id, int t0, int n) {
    float4* tile = reinterpret_cast<float4*>(c.lds);
    for (int k = c.tid; k < n; k += NT) {  // every workgroup stages its own copy
        int j = t0 + k;
        tile[k] = make_float4(pts[j * 3], pts[j * 3 + 1], pts[j * 3 + 2], __int_as_float(id[j]));
    }
    __syncthreads();
}

// One scan of all queries q in [0, nq) (positions qp, ids qid) against all
// elements of the other side (positions ep, ids eid, count ne), for accepted
// pairs within the cache radius.  SYM: partners j > q go to the "up" list, j < q
// to the "lo" list (j == q skipped); otherwise every partner is "up".
// FILL=false adds the counts into up_n[q] / lo_n[q]; FILL=true writes the
// partners, in increasing j, at cursors up_n[q] / lo_n[q] (advanced).  Counts
// and cursors live in registers while a row is scanned.  All threads.
template <bool SYM, bool FILL>
__device__ void pair_scan(Ctx& c, const PairGraph& G, const float* qp, const int* qid, int nq, const float* ep,
                          const int* eid, int ne, int* up_n, int* lo_n, int* up_dst, int* lo_dst) {
    const float4* tile = reinterpret_cast<const float4*>(c.lds);
    const float cc2 = G.cache_cut2;
    for (int t0 = 0; t0 < ne; t0 += TILE) {
        int tn = min(TILE, ne - t0);
        stage_tile(c, ep, eid, t0, tn);
        for (int q = c.gtid; q < nq; q += c.gsize) {
            float x0 = qp[q * 3], x1 = qp[q * 3 + 1], x2 = qp[q * 3 + 2];
            int a = qid[q];
            int up = FILL ? up_n[q] : 0;
            int lo = (FILL && SYM) ? lo_n[q] : 0;
            for (int k = 0; k < tn; ++k) {
                float4 e = tile[k];
                float d2 = refine_dist2(x0 - e.x, x1 - e.y, x2 - e.z);
                int j = t0 + k;
                if (d2 < cc2 && !(SYM && j == q) && pg_accept(G, a, __float_as_int(e.w))) {
                    if (!SYM || j > q) {
                        if (FILL) up_dst[up] = j;
                        ++up;
                    } else {
                        if (FILL) lo_dst[lo] = j;
                        ++lo;
                    }
                }
            }
            if (FILL) {
                up_n[q] = up;
                if (SYM) lo_n[q] = lo;
            } else {
                up_n[q] += up;
                if (SYM) lo_n[q] += lo;
            }
        }
        __syncthreads();  // the tile is restaged next
    }
    team_sync(c);
}

// ---- cache rebuild: cell grid in LDS ----------------------------------------
// The other side is counting-sorted into hashed cells of edge >= the cache
// radius, kept in LDS as (bucket offsets, float4 elements (x, y, z, index)),
// each bucket sorted by element index.  A query visits its 27 neighbouring
// cells in a fixed order, so rows come out in a canonical order.  When the
// grid does not fit in LDS the scan streams the other side in tiles instead
// (same emission semantics, partners in increasing index).
struct LGrid {
    float ox, oy, oz, inv;
    int mask;
    int* ofs;      // [mask+2]
    float4* el;    // [m]  (x, y, z, element index)
    int* eid;      // [m]  element ids, same order (no global loads in the scan)
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
__device__ LGrid lgrid_build(Ctx& c, const PairGraph& G, const float* pts, const int* ids, int m) {
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
    g.ofs = reinterpret_cast<int*>(c.lds);
    g.el = reinterpret_cast<float4*>(c.lds + ((nb + 4) & ~3));
    g.eid = reinterpret_cast<int*>(g.el + m);
    for (int b = c.tid; b <= nb; b += NT) g.ofs[b] = 0;
    __syncthreads();
    for (int i = c.tid; i < m; i += NT) {
        int b = cell_bucket(g.mask, cell_coord(pts[i * 3], g.ox, g.inv), cell_coord(pts[i * 3 + 1], g.oy, g.inv),
                            cell_coord(pts[i * 3 + 2], g.oz, g.inv));
        atomicAdd(&g.ofs[b + 1], 1);
    }
    __syncthreads();
    block_exscan(c, g.ofs + 1, nb);  // ofs[b+1] = start of bucket b
    for (int i = c.tid; i < m; i += NT) {
        float x = pts[i * 3], y = pts[i * 3 + 1], z = pts[i * 3 + 2];
        int b = cell_bucket(g.mask, cell_coord(x, g.ox, g.inv), cell_coord(y, g.oy, g.inv), cell_coord(z, g.oz, g.inv));
        int p = atomicAdd(&g.ofs[b + 1], 1);  // ends at the start of bucket b+1
        g.el[p] = make_float4(x, y, z, __int_as_float(i));
        g.eid[p] = ids[i];
    }
    __syncthreads();
    // canonical order inside every bucket: by element index
    for (int b = c.tid; b < nb; b += NT) {
        int lo = g.ofs[b], hi = g.ofs[b + 1];
        for (int k = lo + 1; k < hi; ++k) {
            float4 v = g.el[k];
            int vid = g.eid[k];
            int key = __float_as_int(v.w);
            int t = k - 1;
            while (t >= lo && __float_as_int(g.el[t].w) > key) {
                g.el[t + 1] = g.el[t];
                g.eid[t + 1] = g.eid[t];
                --t;
            }
            g.el[t + 1] = v;
            g.eid[t + 1] = vid;
        }
    }
    __syncthreads();
    return g;
}

// Query rows against the grid; same contract as pair_scan.
template <bool SYM, bool FILL>
__device__ void pair_scan_grid(Ctx& c, const PairGraph& G, const LGrid& g, const float* qp, const int* qid, int nq,
                               int* up_n, int* lo_n, int* up_dst, int* lo_dst) {
    const float cc2 = G.cache_cut2;
    for (int q = c.gtid; q < nq; q += c.gsize) {
        float x0 = qp[q * 3], x1 = qp[q * 3 + 1], x2 = qp[q * 3 + 2];
        int a = qid[q];
        int up = FILL ? up_n[q] : 0;
        int lo = (FILL && SYM) ? lo_n[q] : 0;
        int cx = cell_coord(x0, g.ox, g.inv), cy = cell_coord(x1, g.oy, g.inv), cz = cell_coord(x2, g.oz, g.inv);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int ax = cx + dx, ay = cy + dy, az = cz + dz;
                    int b = cell_bucket(g.mask, ax, ay, az);
                    int e1 = g.ofs[b + 1];
                    for (int t = g.ofs[b]; t < e1; ++t) {
                        float4 e = g.el[t];
                        float d2 = refine_dist2(x0 - e.x, x1 - e.y, x2 - e.z);
                        if (!(d2 < cc2)) continue;
                        if (cell_coord(e.x, g.ox, g.inv) != ax || cell_coord(e.y, g.oy, g.inv) != ay ||
                            cell_coord(e.z, g.oz, g.inv) != az)
                            continue;  // another cell sharing the bucket
                        int j = __float_as_int(e.w);
                        if (SYM && j == q) continue;
                        if (!pg_accept(G, a, g.eid[t])) continue;
                        if (!SYM || j > q) {
                            if (FILL) up_dst[up] = j;
                            ++up;
                        } else {
                            if (FILL) lo_dst[lo] = j;
                            ++lo;
                        }
                    }
                }
        if (FILL) {
            up_n[q] = up;
            if (SYM) lo_n[q] = lo;
        } else {
            up_n[q] += up;
            if (SYM) lo_n[q] += lo;
        }
    }
    team_sync(c);
}

// Scan by grid when the other side fits in LDS, else by tiles.
template <bool SYM, bool FILL>
__device__ void pair_scan_any(Ctx& c, const PairGraph& G, const float* qp, const int* qid, int nq, const float* ep,
                              const int* eid, int ne, int* up_n, int* lo_n, int* up_dst, int* lo_dst) {
    if (lgrid_fits(ne)) {
        unsigned long long t0 = ptick(c);
        LGrid g = lgrid_build(c, G, ep, eid, ne);
        pacc(c, PROF_RB_GRID, t0);
        t0 = ptick(c);
        pair_scan_grid<SYM, FILL>(c, G, g, qp, qid, nq, up_n, lo_n, up_dst, lo_dst);
        pacc(c, PROF_RB_SCAN, t0);
    } else {
        pair_scan<SYM, FILL>(c, G, qp, qid, nq, ep, eid, ne, up_n, lo_n, up_dst, lo_dst);
    }
}

// Slot lookup table: open addressing on key i*n2 + j (HCAP a power of two).
__device__ __forceinline__ uint32_t slot_hash(uint32_t k) { return k * 2654435761u; }

// Rebuild rows and transposed rows from pos_pack.  Sets state[1] on overflow
// (the host grows the capacity and repeats the run).
//   forward rows:    side-1 element i, partners j (symmetric: j > i);
//   transposed rows: side-2 element j, partners i (symmetric: i < j), stored
//                    as the forward slot of (i, j) found through a hash table.
// Both row orders are canonical (see the scans above).
__device__ void pg_rebuild(Ctx& c, const PairGraph& G) {
    const int* id1 = c.CI + G.id1;
    const int* id2 = c.CI + G.id2;
    float* pack = c.A + G.pos_pack;
    float* cache = c.A + G.cache_pos;
    const float* p1 = pack;
    const float* p2 = pg_side2(c, G);
    int* off = c.I + G.off;
    int* nbr = c.I + G.nbr;
    int* row = c.I + G.row;
    int* toff = c.I + G.toff;
    int* tslot = c.I + G.tslot;
    int* tmp = c.I + G.tmp;
    int* hkey = c.I + G.hkey;
    int* hval = c.I + G.hval;
    const bool sym = G.sym != 0;
    int npack = sym ? G.n1 : G.n1 + G.n2;
    for (int i = c.gtid; i < npack * 3; i += c.gsize) cache[i] = pack[i];
    for (int i = c.gtid; i <= G.n1; i += c.gsize) off[i] = 0;
    for (int j = c.gtid; j <= G.n2; j += c.gsize) toff[j] = 0;
    team_sync(c);
    unsigned long long t0 = ptick(c);
    // counts (each query row is owned by one thread)
    if (sym) {
        pair_scan_any<true, false>(c, G, p1, id1, G.n1, p1, id1, G.n1, off, toff, nullptr, nullptr);
    } else {
        pair_scan_any<false, false>(c, G, p1, id1, G.n1, p2, id2, G.n2, off, nullptr, nullptr, nullptr);
        pair_scan_any<false, false>(c, G, p2, id2, G.n2, p1, id1, G.n1, toff, nullptr, nullptr, nullptr);
    }
    int E = team_exscan(c, off, G.n1 + 1);
    int Et = team_exscan(c, toff, G.n2 + 1);
    if (E > G.ecap || Et != E) {  // Et == E always holds: both scans apply one symmetric test
        if (c.gtid == 0) { c.I[G.state + 1] = E > G.ecap ? 1 : 2; c.I[G.state + 0] = 0; }
        team_sync(c);
        return;
    }
    pacc(c, PROF_RB_COUNT, t0);
    t0 = ptick(c);
    // fill: forward partners into nbr, transposed partners into tmp
    for (int i = c.gtid; i < G.n1; i += c.gsize) row[i] = off[i];        // row[] as a cursor for now
    for (int j = c.gtid; j < G.n2; j += c.gsize) tslot[j] = toff[j];     // tslot[] as a cursor for now
    for (int h = c.gtid; h < G.hcap; h += c.gsize) hkey[h] = -1;
    team_sync(c);
    if (sym) {
        pair_scan_any<true, true>(c, G, p1, id1, G.n1, p1, id1, G.n1, row, tslot, nbr, tmp);
    } else {
        pair_scan_any<false, true>(c, G, p1, id1, G.n1, p2, id2, G.n2, row, nullptr, nbr, nullptr);
        pair_scan_any<false, true>(c, G, p2, id2, G.n2, p1, id1, G.n1, tslot, nullptr, tmp, nullptr);
    }
    for (int i = c.gtid; i < G.n1; i += c.gsize)
        for (int s = off[i]; s < off[i + 1]; ++s) row[s] = i;
    team_sync(c);
    pacc(c, PROF_RB_FILL, t0);
    t0 = ptick(c);
    // slot table: (i, j) -> forward slot
    const uint32_t hm = (uint32_t)G.hcap - 1u;
    const int n2 = G.n2;
    int bad = 0;
    for (int s = c.gtid; s < E; s += c.gsize) {
        int key = row[s] * n2 + nbr[s];
        uint32_t h = slot_hash((uint32_t)key) & hm;
        for (int probe = 0; probe < G.hcap; ++probe) {
            int prev = atomicCAS(&hkey[h], -1, key);
            if (prev == -1) { hval[h] = s; break; }
            h = (h + 1u) & hm;
        }
    }
    team_sync(c);
    // transposed entries -> forward slots
    int nt = sym ? G.n1 : G.n2;
    for (int j = c.gtid; j < nt; j += c.gsize)
        for (int t = toff[j]; t < toff[j + 1]; ++t) {
            int key = tmp[t] * n2 + j;
            uint32_t h = slot_hash((uint32_t)key) & hm;
            int found = -1;
            for (int probe = 0; probe < G.hcap; ++probe) {
                int k = hkey[h];
                if (k == key) { found = hval[h]; break; }
                if (k == -1) break;
                h = (h + 1u) & hm;
            }
            if (found < 0) { bad = 1; found = 0; }
            tslot[t] = found;
        }
    int anybad = team_or(c, bad);
    pacc(c, PROF_RB_MAP, t0);
    if (c.gtid == 0) {
        c.I[G.state + 0] = anybad ? 0 : 1;
        if (anybad) c.I[G.state + 1] = 2;
        c.I[G.state + 2] += 1;
        c.I[G.state + 3] = E;
    }
    team_sync(c);
}

// ---- edge functions -----------------------------------------------------
// quadspline (src/bead_interaction.h:31-85): radial clamped splines times
// angular splines of the two bead directions.
template <int NKA, int NK>
__device__ __forceinline__ float quadspline(const float* p, const float* x1, const float* x2,
                                            float inv_dtheta, float inv_dx, float* d1, float* d2) {
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
__device__ __forceinline__ float protein_hbond_edge(const float* p, const float* x1, const float* x2,
                                                    float* d1, float* d2) {
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
template <bool WEIGHTED>
__device__ __forceinline__ float env_edge(const float* p, const float* x1, const float* x2, float* d1, float* d2) {
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
    const float* aff = c.A + op.in_out[0];
    int w = op.in_w[0];
    const float* ref = c.CF + op.ia[1];
    const int* nat = c.CI + op.ia[2];
    const int* res = c.CI + op.ia[0];
    int R1 = res[r1], R2 = res[r2];
    const float* a1 = aff + R1 * w;
    const float* a2 = aff + R2 * w;
    float U1[9], U2[9];
    quat_to_rot(a1 + 3, U1);
    quat_to_rot(a2 + 3, U2);
    float X1[4][3], X2[4][3];
    int n1 = nat[r1], n2 = nat[r2];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float* q = ref + (r1 * 4 + a) * 3;
        const float* s = ref + (r2 * 4 + a) * 3;
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

// Stage the graph's coefficient table in LDS (when it has one that fits).
__device__ const float* pg_stage_params(Ctx& c, const PairGraph& G) {
    const float* src = c.CF + G.param;
    if (G.param_len <= 0) return src;
    for (int k = c.tid; k < G.param_len; k += NT) c.lds[k] = src[k];  // per workgroup
    __syncthreads();
    return c.lds;
}

// Refine every cached slot against the cutoff and evaluate the active edges.
template <int KIND, int DD1, int DD2>
__device__ void pg_edges_t(Ctx& c, const PairGraph& G, const OpRec& op) {
    const float* param = pg_stage_params(c, G);
    const int* nbr = c.I + G.nbr;
    const int* row = c.I + G.row;
    int* act = c.I + G.act;
    float* val = c.A + G.val;
    float* D1 = c.A + G.d1;
    float* D2 = c.A + G.d2;
    const float* pack = c.A + G.pos_pack;
    const float* pk2 = pg_side2(c, G);
    const int* loc1 = c.CI + G.loc1;
    const int* loc2 = c.CI + G.loc2;
    const int* t1 = c.CI + G.type1;
    const int* t2 = c.CI + G.type2;
    const float* P1 = c.A + G.out1;
    const float* P2 = c.A + G.out2;
    const int E = c.I[G.state + 3];
    const int cap = G.ecap;
    for (int s = c.gtid; s < E; s += c.gsize) {
        int i1 = row[s], j = nbr[s];
        const float* xp = pack + i1 * 3;
        const float* yp = pk2 + j * 3;
        float d2 = refine_dist2(xp[0] - yp[0], xp[1] - yp[1], xp[2] - yp[2]);
        int a = d2 < G.cutoff2;
        act[s] = a;
        if (!a) { val[s] = 0.f; continue; }
        const float* x1 = P1 + loc1[i1] * G.w1;
        const float* x2 = P2 + loc2[j] * G.w2;
        const float* p = param + (t1[i1] * G.n_type2 + t2[j]) * G.n_param;
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

template <int D1N, int D2N>
__device__ void pg_gather_t(Ctx& c, const PairGraph& G) {
    const int* off = c.I + G.off;
    const int* toff = c.I + G.toff;
    const int* tslot = c.I + G.tslot;
    const int* act = c.I + G.act;
    const float* es = c.A + G.esens;
    const float* D1 = c.A + G.d1;
    const float* D2 = c.A + G.d2;
    const int* loc1 = c.CI + G.loc1;
    const int* loc2 = c.CI + G.loc2;
    float* S1 = c.A + G.sens1;
    float* S2 = c.A + G.sens2;
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
            float* dst = S1 + loc1[i] * G.w1;
#pragma unroll
            for (int q = 0; q < D1N; ++q) dst[q] += acc[q];
        }
    }
    team_sync(c);
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
            float* dst = S2 + loc2[j] * G.w2;
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
__device__ void pg_row_sums(Ctx& c, const PairGraph& G, bool side2, float* out, int stride) {
    const int* off = c.I + (side2 ? G.toff : G.off);
    const int* tslot = c.I + G.tslot;
    const int* act = c.I + G.act;
    const float* val = c.A + G.val;
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
