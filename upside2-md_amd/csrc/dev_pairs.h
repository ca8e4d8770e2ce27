// Pair graphs on the device: Verlet cache, exact edge refine, per-edge spline
// energies and the fixed-order sensitivity gather.
//
// Reference: PairlistComputation / InteractionGraph (src/interaction_graph.h).
// The edge set of a graph is {(i1,i2) acceptable : dist2 < cutoff^2} where
// dist2 is computed from the current packed coordinates exactly as the
// reference build evaluates it, fma(x,x, fma(z,z, y*y)) (the Float4 sum of
// src/interaction_graph.h:239 contracted by the compiler).  The cache (radius
// cutoff + buffer, rebuilt when any element drifts more than buffer/2,
// src/interaction_graph.h:51-169) only bounds which pairs are examined.
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

__device__ __forceinline__ const float* pg_p2(const Ctx& c, const PairGraph& G, int j) {
    return c.A + G.pos_pack + (G.sym ? j : G.n1 + j) * 3;
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
    for (int i = c.tid; i < n; i += NT) {
        const float* src = (i < G.n1) ? P1 + loc1[i] * G.w1 : P2 + loc2[i - G.n1] * G.w2;
        float x = src[0], y = src[1], z = src[2];
        pack[i * 3 + 0] = x; pack[i * 3 + 1] = y; pack[i * 3 + 2] = z;
        float dx = x - cache[i * 3 + 0], dy = y - cache[i * 3 + 1], dz = z - cache[i * 3 + 2];
        if (dx * dx + dy * dy + dz * dz > G.drift2) moved = 1;
    }
    int any = __syncthreads_or(moved | (valid ? 0 : 1));
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

// Rebuild the candidate rows and the transposed lists from pos_pack.
__device__ void pg_rebuild(Ctx& c, const PairGraph& G) {
    const int* id1 = c.CI + G.id1;
    const int* id2 = c.CI + G.id2;
    float* pack = c.A + G.pos_pack;
    float* cache = c.A + G.cache_pos;
    int* cnt = c.I + G.cnt;
    int* nbr = c.I + G.nbr;
    int* tcnt = c.I + G.tcnt;
    int* tslot = c.I + G.tslot;
    int n = G.sym ? G.n1 : G.n1 + G.n2;
    for (int i = c.tid; i < n * 3; i += NT) cache[i] = pack[i];
    int overflow = 0;
    for (int i1 = c.tid; i1 < G.n1; i1 += NT) {
        const float* x = pack + i1 * 3;
        float x0 = x[0], x1 = x[1], x2 = x[2];
        int a = id1[i1];
        int k = 0;
        for (int j = G.sym ? i1 + 1 : 0; j < G.n2; ++j) {
            const float* y = pg_p2(c, G, j);
            float dx = x0 - y[0], dy = x1 - y[1], dz = x2 - y[2];
            if (refine_dist2(dx, dy, dz) < G.cache_cut2 && pg_accept(G, a, id2[j])) {
                if (k < G.maxn) nbr[i1 * G.maxn + k] = j;
                ++k;
            }
        }
        if (k > G.maxn) { overflow = 1; k = G.maxn; }
        cnt[i1] = k;
    }
    __syncthreads();
    for (int j = c.tid; j < G.n2; j += NT) {
        const float* y = pg_p2(c, G, j);
        float y0 = y[0], y1 = y[1], y2 = y[2];
        int b = id2[j];
        int t = 0;
        int lim = G.sym ? j : G.n1;
        for (int i1 = 0; i1 < lim; ++i1) {
            const float* x = pack + i1 * 3;
            float dx = x[0] - y0, dy = x[1] - y1, dz = x[2] - y2;
            if (refine_dist2(dx, dy, dz) < G.cache_cut2 && pg_accept(G, id1[i1], b)) {
                int k = lower_bound_i(nbr + i1 * G.maxn, cnt[i1], j);
                if (k < cnt[i1] && nbr[i1 * G.maxn + k] == j) {
                    if (t < G.maxt) tslot[j * G.maxt + t] = i1 * G.maxn + k;
                    ++t;
                }
            }
        }
        if (t > G.maxt) { overflow = 1; t = G.maxt; }
        tcnt[j] = t;
    }
    int of = __syncthreads_or(overflow);
    if (c.tid == 0) {
        c.I[G.state + 0] = 1;
        if (of) c.I[G.state + 1] = 1;
        c.I[G.state + 2] += 1;
    }
    __syncthreads();
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
        for (int k = 0; k < 3; ++k) {
            drOC[k] = c1 * r[k];
            drHN[k] = c2 * r[k];
            dH[k] = c0 * r[k] + (c1 * inv) * (x2[3 + k] - dHOC * r[k]) + (c2 * inv) * (x1[3 + k] + dOHN * r[k]);
        }
    }
    float hb_log = (1.f <= hb) ? 100.f : -logf(1.f - hb);
    float pref = fminf(1.f / (1.f - hb), 1e5f);
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
// repulsion between the placed backbone atoms of two residues; derivatives are
// (force, torque) on each residue frame.
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
    for (int a = 0; a < n1; ++a) {
        const float* q = ref + (r1 * 4 + a) * 3;
        for (int k = 0; k < 3; ++k) X1[a][k] = U1[3 * k] * q[0] + U1[3 * k + 1] * q[1] + U1[3 * k + 2] * q[2] + a1[k];
    }
    for (int a = 0; a < n2; ++a) {
        const float* q = ref + (r2 * 4 + a) * 3;
        for (int k = 0; k < 3; ++k) X2[a][k] = U2[3 * k] * q[0] + U2[3 * k + 1] * q[1] + U2[3 * k + 2] * q[2] + a2[k];
    }
    const float cut2 = 3.f * 3.f + 0.1f * 3.f;
    const float sharp = 1.f / (3.0f * 0.10f);
    float pot = 0.f;
    for (int k = 0; k < 6; ++k) { d1[k] = 0.f; d2[k] = 0.f; }
    for (int i = 0; i < n1; ++i) {
        for (int j = 0; j < n2; ++j) {
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

// Refine every cached slot against the cutoff and evaluate active edges.
__device__ void pg_edges(Ctx& c, const PairGraph& G, const OpRec& op) {
    const int* cnt = c.I + G.cnt;
    const int* nbr = c.I + G.nbr;
    int* act = c.I + G.act;
    float* val = c.A + G.val;
    float* D1 = c.A + G.d1;
    float* D2 = c.A + G.d2;
    const float* pack = c.A + G.pos_pack;
    const int* loc1 = c.CI + G.loc1;
    const int* loc2 = c.CI + G.loc2;
    const int* t1 = c.CI + G.type1;
    const int* t2 = c.CI + G.type2;
    const float* P1 = c.A + G.out1;
    const float* P2 = c.A + G.out2;
    const float* param = c.CF + G.param;
    int wave = c.tid >> 6, lane = c.tid & 63;
    for (int i1 = wave; i1 < G.n1; i1 += NWAVE) {
        int n = cnt[i1];
        const float* xp = pack + i1 * 3;
        const float* x1 = P1 + loc1[i1] * G.w1;
        for (int k = lane; k < n; k += 64) {
            int slot = i1 * G.maxn + k;
            int j = nbr[slot];
            const float* yp = pg_p2(c, G, j);
            float d2 = refine_dist2(xp[0] - yp[0], xp[1] - yp[1], xp[2] - yp[2]);
            int a = d2 < G.cutoff2;
            act[slot] = a;
            if (!a) { val[slot] = 0.f; continue; }
            const float* x2 = P2 + loc2[j] * G.w2;
            const float* p = param + (t1[i1] * G.n_type2 + t2[j]) * G.n_param;
            float e1[7], e2[7];
            float v = 0.f;
            switch (G.kind) {
            case EDGE_ROTAMER:
                v = quadspline<15, 12>(p, x1, x2, 6.f, 1.f / 0.7f, e1, e2);
                break;
            case EDGE_HBOND_COVERAGE: {
                float cov = quadspline<15, 10>(p, x1, x2, 6.f, 1.f / 0.625f, e1, e2);
                float om = 1.f - x1[6];
                float pref = om * om;
                for (int q = 0; q < 6; ++q) { e1[q] *= pref; e2[q] *= pref; }
                e1[6] = -cov * om * 2.f;
                v = pref * cov;
            } break;
            case EDGE_PROTEIN_HBOND:
                v = protein_hbond_edge(p, x1, x2, e1, e2);
                break;
            case EDGE_ENV_WEIGHTED:
                v = env_edge<true>(p, x1, x2, e1, e2);
                break;
            case EDGE_ENV_HBOND:
                v = env_edge<false>(p, x1, x2, e1, e2);
                break;
            case EDGE_BACKBONE:
                v = backbone_edge(c, op, i1, j, e1, e2);
                break;
            }
            val[slot] = v;
            for (int q = 0; q < G.ddim1; ++q) D1[(size_t)slot * G.ddim1 + q] = e1[q];
            for (int q = 0; q < G.ddim2; ++q) D2[(size_t)slot * G.ddim2 + q] = e2[q];
        }
    }
    __syncthreads();
}

// Gather esens*deriv into the producers in a fixed order.  `col0` offsets the
// target columns (backbone pairs write into the 6 leading affine columns).
__device__ void pg_gather(Ctx& c, const PairGraph& G) {
    const int* cnt = c.I + G.cnt;
    const int* tcnt = c.I + G.tcnt;
    const int* tslot = c.I + G.tslot;
    const int* act = c.I + G.act;
    const float* es = c.A + G.esens;
    const float* D1 = c.A + G.d1;
    const float* D2 = c.A + G.d2;
    const int* loc1 = c.CI + G.loc1;
    const int* loc2 = c.CI + G.loc2;
    float* S1 = c.A + G.sens1;
    float* S2 = c.A + G.sens2;
    for (int i = c.tid; i < G.n1; i += NT) {
        float acc[7] = {0, 0, 0, 0, 0, 0, 0};
        int n = cnt[i];
        for (int k = 0; k < n; ++k) {
            int slot = i * G.maxn + k;
            if (!act[slot]) continue;
            float s = es[slot];
            for (int q = 0; q < G.ddim1; ++q) acc[q] += s * D1[(size_t)slot * G.ddim1 + q];
        }
        if (G.sym) {
            int m = tcnt[i];
            for (int t = 0; t < m; ++t) {
                int slot = tslot[i * G.maxt + t];
                if (!act[slot]) continue;
                float s = es[slot];
                for (int q = 0; q < G.ddim2; ++q) acc[q] += s * D2[(size_t)slot * G.ddim2 + q];
            }
        }
        float* dst = S1 + loc1[i] * G.w1;
        for (int q = 0; q < G.ddim1; ++q) dst[q] += acc[q];
    }
    __syncthreads();
    if (G.sym) return;
    for (int j = c.tid; j < G.n2; j += NT) {
        float acc[7] = {0, 0, 0, 0, 0, 0, 0};
        int m = tcnt[j];
        for (int t = 0; t < m; ++t) {
            int slot = tslot[j * G.maxt + t];
            if (!act[slot]) continue;
            float s = es[slot];
            for (int q = 0; q < G.ddim2; ++q) acc[q] += s * D2[(size_t)slot * G.ddim2 + q];
        }
        float* dst = S2 + loc2[j] * G.w2;
        for (int q = 0; q < G.ddim2; ++q) dst[q] += acc[q];
    }
    __syncthreads();
}

// Prepare a graph for an evaluation: pack, maybe rebuild, refine and evaluate.
__device__ void pg_refresh(Ctx& c, const PairGraph& G, const OpRec& op) {
    if (pg_pack_and_check(c, G)) pg_rebuild(c, G);
    pg_edges(c, G, op);
}

}  // namespace upside
