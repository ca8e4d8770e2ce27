// Rotamer side-chain free energy by loopy belief propagation on the device
// (src/rotamer.cpp:628-1169), one workgroup per system.
//
// Residue nodes carry 1, 3 or 6 rotamer states.  Residue pairs are candidate
// "edges" kept per cache rebuild (any bead pair within the cache radius); an
// edge is present in an evaluation when at least one of its bead pairs is in
// the refined edge set, which is the reference's EdgeLocator edge set
// (src/rotamer.cpp:134-206,383-396).  Messages are updated synchronously
// (Jacobi) per edge, node beliefs are products over incident edges, with the
// reference's schedule: an undamped first pass whose node beliefs are swapped
// out, then damped passes in chunks of `iteration_chunk_size`, stopping when
// the one-sided max change is below tol (src/rotamer.cpp:1089-1162).
#pragma once
#include "dev_pairs.h"

namespace upside {

struct RotView {
    const int* nrot; const int* bead0; const int* brot;
    int* ccnt; int* cnbr; int* icnt; int* iinc; int* cpres;
    float* cP; float* cmsg; float* cmarg; float* nprob; float* ncur; float* nold; float* noff;
    int maxc, maxi;
};

__device__ RotView rot_view(Ctx& c, const RotamerDev& R) {
    RotView v;
    v.nrot = c.CI + R.node_nrot; v.bead0 = c.CI + R.node_bead0; v.brot = c.CI + R.bead_rot;
    v.ccnt = c.I + R.ccnt; v.cnbr = c.I + R.cnbr; v.icnt = c.I + R.icnt; v.iinc = c.I + R.iinc;
    v.cpres = c.I + R.cpres;
    v.cP = c.A + R.cP; v.cmsg = c.A + R.cmsg; v.cmarg = c.A + R.cmarg;
    v.nprob = c.A + R.nprob; v.ncur = c.A + R.ncur; v.nold = c.A + R.nold; v.noff = c.A + R.noff;
    v.maxc = R.maxc; v.maxi = R.maxi;
    return v;
}

// Node-pair candidates from the bead cache positions (called after a bead
// graph rebuild).  Row A lists partner nodes B>A in increasing order; the
// incidence list of node X names (candidate, side) for A<X first, then its own.
__device__ void rot_rebuild(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const float* cache = c.A + G.cache_pos;
    int overflow = 0;
    for (int a = c.tid; a < R.n_node; a += NT) {
        int k = 0;
        for (int b = a + 1; b < R.n_node; ++b) {
            bool hit = false;
            for (int i = v.bead0[a]; i < v.bead0[a + 1] && !hit; ++i)
                for (int j = v.bead0[b]; j < v.bead0[b + 1]; ++j) {
                    float dx = cache[i * 3] - cache[j * 3], dy = cache[i * 3 + 1] - cache[j * 3 + 1],
                          dz = cache[i * 3 + 2] - cache[j * 3 + 2];
                    if (refine_dist2(dx, dy, dz) < R.edge_cut2) { hit = true; break; }
                }
            if (hit) {
                if (k < v.maxc) v.cnbr[a * v.maxc + k] = b;
                ++k;
            }
        }
        if (k > v.maxc) { overflow = 1; k = v.maxc; }
        v.ccnt[a] = k;
    }
    __syncthreads();
    for (int x = c.tid; x < R.n_node; x += NT) {
        int t = 0;
        for (int a = 0; a < x; ++a) {
            int k = lower_bound_i(v.cnbr + a * v.maxc, v.ccnt[a], x);
            if (k < v.ccnt[a] && v.cnbr[a * v.maxc + k] == x) {
                if (t < v.maxi) v.iinc[x * v.maxi + t] = (a * v.maxc + k) * 2 + 1;
                ++t;
            }
        }
        for (int k = 0; k < v.ccnt[x]; ++k) {
            if (t < v.maxi) v.iinc[x * v.maxi + t] = (x * v.maxc + k) * 2 + 0;
            ++t;
        }
        if (t > v.maxi) { overflow = 1; t = v.maxi; }
        v.icnt[x] = t;
    }
    int of = __syncthreads_or(overflow);
    if (c.tid == 0 && of) c.I[G.state + 1] = 1;
    __syncthreads();
}

// Orientation of candidate (a,b): rows of P belong to the node with fewer
// rotamers (src/rotamer.cpp:898); for equal counts the lower node comes first.
__device__ __forceinline__ bool rot_swapped(const RotView& v, int a, int b) { return v.nrot[a] > v.nrot[b]; }

// Fill node probabilities and candidate edge matrices for this evaluation.
__device__ void rot_fill(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int* loc1 = c.CI + G.loc1;
    const int* cnt = c.I + G.cnt;
    const int* nbr = c.I + G.nbr;
    const int* act = c.I + G.act;
    const float* val = c.A + G.val;
    // node energies -> probabilities (fill_holders, convert_energy_to_prob, :239-256,867-880)
    for (int n = c.tid; n < R.n_node; n += NT) {
        float e[6] = {0, 0, 0, 0, 0, 0};
        for (int b = v.bead0[n]; b < v.bead0[n + 1]; ++b) {
            float s = 0.f;
            for (int q = 0; q < R.n_prob; ++q) s += c.A[R.prob_out[q] + loc1[b] * R.prob_w[q]];
            e[v.brot[b]] += s;
        }
        int nr = v.nrot[n];
        float mn = e[0];
        for (int r = 1; r < nr; ++r) mn = fminf(mn, e[r]);
        v.noff[n] = mn;
        for (int r = 0; r < 6; ++r) v.nprob[n * 6 + r] = r < nr ? expf(mn - e[r]) : 0.f;
    }
    // candidate edge probabilities P[r1][r2] = prod exp(-E) over active bead pairs
    for (int a = c.tid; a < R.n_node; a += NT) {
        for (int k = 0; k < v.ccnt[a]; ++k) {
            int ci = a * v.maxc + k;
            int b = v.cnbr[ci];
            bool sw = rot_swapped(v, a, b);
            int n1 = sw ? v.nrot[b] : v.nrot[a], n2 = sw ? v.nrot[a] : v.nrot[b];
            float* P = v.cP + (size_t)ci * 36;
            for (int q = 0; q < 36; ++q) P[q] = ((q / 6) < n1 && (q % 6) < n2) ? 1.f : 0.f;
            int pres = 0;
            int jb0 = v.bead0[b], jb1 = v.bead0[b + 1];
            for (int i = v.bead0[a]; i < v.bead0[a + 1]; ++i) {
                const int* row = nbr + i * G.maxn;
                int k0 = lower_bound_i(row, cnt[i], jb0);
                for (int s = k0; s < cnt[i] && row[s] < jb1; ++s) {
                    int slot = i * G.maxn + s;
                    if (!act[slot]) continue;
                    pres = 1;
                    int ra = v.brot[i], rb = v.brot[row[s]];
                    int r1 = sw ? rb : ra, r2 = sw ? ra : rb;
                    P[r1 * 6 + r2] *= expf(-val[slot]);
                }
            }
            v.cpres[ci] = pres;
        }
    }
    __syncthreads();
    // fold present 1-rotamer edges into the partner's probabilities (:398-405)
    for (int x = c.tid; x < R.n_node; x += NT) {
        if (v.nrot[x] == 1) continue;
        for (int t = 0; t < v.icnt[x]; ++t) {
            int e = v.iinc[x * v.maxi + t];
            int ci = e >> 1;
            if (!v.cpres[ci]) continue;
            int a = ci / v.maxc, b = v.cnbr[ci];
            int other = (e & 1) ? a : b;
            if (v.nrot[other] != 1) continue;
            const float* P = v.cP + (size_t)ci * 36;  // row 0 belongs to the 1-rotamer node
            for (int r = 0; r < v.nrot[x]; ++r) v.nprob[x * 6 + r] *= P[r];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool rot_bp_edge(const RotView& v, int ci, int& f, int& s) {
    // BP runs on present edges between 3- or 6-rotamer nodes
    int a = ci / v.maxc, b = v.cnbr[ci];
    if (v.nrot[a] == 1 || v.nrot[b] == 1) return false;
    bool sw = rot_swapped(v, a, b);
    f = sw ? b : a; s = sw ? a : b;
    return true;
}

// One synchronous message pass: new messages from old node beliefs `nold_`
// and old messages (half `po` of cmsg); each new message is scaled to unit L1
// norm (src/rotamer.cpp:515-568).
__device__ void rot_edge_pass(Ctx& c, const RotamerDev& R, const float* nold_, int po) {
    RotView v = rot_view(c, R);
    int total = R.n_node * v.maxc;
    for (int ci = c.tid; ci < total; ci += NT) {
        int a = ci / v.maxc, k = ci % v.maxc;
        if (k >= v.ccnt[a] || !v.cpres[ci]) continue;
        int f, s;
        if (!rot_bp_edge(v, ci, f, s)) continue;
        int n1 = v.nrot[f], n2 = v.nrot[s];
        const float* P = v.cP + (size_t)ci * 36;
        const float* mo = v.cmsg + (size_t)ci * 24 + po * 12;
        float* mn = v.cmsg + (size_t)ci * 24 + (1 - po) * 12;
        float v1[6], v2[6];
        for (int r = 0; r < n1; ++r) v1[r] = nold_[f * 6 + r] / (1e-10f + mo[r]);
        for (int r = 0; r < n2; ++r) v2[r] = nold_[s * 6 + r] / (1e-10f + mo[6 + r]);
        float m1[6], m2[6], s1 = 0.f, s2 = 0.f;
        for (int i = 0; i < n1; ++i) {
            float acc = 0.f;
            for (int j = 0; j < n2; ++j) acc += P[i * 6 + j] * v2[j];
            m1[i] = acc; s1 += acc;
        }
        for (int j = 0; j < n2; ++j) {
            float acc = 0.f;
            for (int i = 0; i < n1; ++i) acc += P[i * 6 + j] * v1[i];
            m2[j] = acc; s2 += acc;
        }
        float i1 = 1.f / s1, i2 = 1.f / s2;
        for (int r = 0; r < 6; ++r) {
            mn[r] = r < n1 ? m1[r] * i1 : 0.f;
            mn[6 + r] = r < n2 ? m2[r] * i2 : 0.f;
        }
    }
    __syncthreads();
}

// Node pass: cur = prob * prod(new messages), renormalised after every factor;
// then cur = (1-d) cur/max(cur) + d*old (standardize_belief_update, :258-273).
// With swap_first, the product goes to `nold_` and cur restarts from prob
// (the first, undamped pass of solve_for_marginals).
__device__ void rot_node_pass(Ctx& c, const RotamerDev& R, float* ncur_, float* nold_, int pn, float damping,
                              bool swap_first) {
    RotView v = rot_view(c, R);
    for (int x = c.tid; x < R.n_node; x += NT) {
        int nr = v.nrot[x];
        if (nr == 1) continue;
        float b[6];
        for (int r = 0; r < 6; ++r) b[r] = v.nprob[x * 6 + r];
        for (int t = 0; t < v.icnt[x]; ++t) {
            int e = v.iinc[x * v.maxi + t];
            int ci = e >> 1;
            if (!v.cpres[ci]) continue;
            int f, s;
            if (!rot_bp_edge(v, ci, f, s)) continue;
            const float* m = v.cmsg + (size_t)ci * 24 + pn * 12 + (f == x ? 0 : 6);
            float sum = 0.f;
            for (int r = 0; r < nr; ++r) { b[r] *= m[r]; sum += b[r]; }
            float inv = 1.f / sum;
            for (int r = 0; r < nr; ++r) b[r] *= inv;
        }
        if (swap_first) {
            float mx = v.nprob[x * 6];
            for (int r = 1; r < nr; ++r) mx = fmaxf(mx, v.nprob[x * 6 + r]);
            for (int r = 0; r < 6; ++r) {
                nold_[x * 6 + r] = b[r];
                ncur_[x * 6 + r] = r < nr ? v.nprob[x * 6 + r] / mx : 0.f;
            }
        } else {
            float mx = b[0];
            for (int r = 1; r < nr; ++r) mx = fmaxf(mx, b[r]);
            float im = 1.f / mx;
            for (int r = 0; r < 6; ++r)
                ncur_[x * 6 + r] = r < nr ? (1.f - damping) * im * b[r] + damping * nold_[x * 6 + r] : 0.f;
        }
    }
    __syncthreads();
}

// Solve; returns the iteration count (the reference's `iter`).
__device__ int rot_solve(Ctx& c, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    int total = R.n_node * v.maxc;
    // initial messages = 1, old beliefs = prob (:1110-1133)
    for (int ci = c.tid; ci < total; ci += NT) {
        float* m = v.cmsg + (size_t)ci * 24;
        for (int r = 0; r < 12; ++r) m[r] = 1.f;
    }
    for (int i = c.tid; i < R.n_node * 6; i += NT) v.nold[i] = v.nprob[i];
    __syncthreads();
    float* cur = v.ncur;
    float* old = v.nold;
    int po = 0;
    rot_edge_pass(c, R, old, po);
    po = 1 - po;  // new messages now live in half po
    rot_node_pass(c, R, cur, old, po, 0.f, true);
    float dev = 1e10f;
    int it = 0;
    for (; dev > R.tol && it < R.max_iter; it += R.chunk) {
        for (int j = 0; j < R.chunk; ++j) {
            float* t = cur; cur = old; old = t;     // swap node beliefs
            rot_edge_pass(c, R, old, po);           // messages old = half po
            po = 1 - po;
            rot_node_pass(c, R, cur, old, po, R.damping, false);
        }
        float d = 0.f;
        for (int x = c.tid; x < R.n_node; x += NT) {
            int nr = v.nrot[x];
            if (nr == 1) continue;
            for (int r = 0; r < nr; ++r) d = fmaxf(d, cur[x * 6 + r] - old[x * 6 + r]);
        }
        dev = block_max(c, d);
    }
    // marginals (:283-290, :425-449); final beliefs may sit in either buffer
    for (int x = c.tid; x < R.n_node; x += NT) {
        int nr = v.nrot[x];
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s += cur[x * 6 + r];
        for (int r = 0; r < 6; ++r)
            v.ncur[x * 6 + r] = nr == 1 ? (r == 0 ? 1.f : 0.f) : (r < nr ? cur[x * 6 + r] / s : 0.f);
    }
    __syncthreads();
    for (int ci = c.tid; ci < total; ci += NT) {
        int a = ci / v.maxc, k = ci % v.maxc;
        if (k >= v.ccnt[a] || !v.cpres[ci]) continue;
        int f, s;
        if (!rot_bp_edge(v, ci, f, s)) continue;
        int n1 = v.nrot[f], n2 = v.nrot[s];
        const float* P = v.cP + (size_t)ci * 36;
        const float* m = v.cmsg + (size_t)ci * 24 + po * 12;
        float b1[6], b2[6];
        for (int r = 0; r < n1; ++r) b1[r] = v.ncur[f * 6 + r] / (1e-10f + m[r]);
        for (int r = 0; r < n2; ++r) b2[r] = v.ncur[s * 6 + r] / (1e-10f + m[6 + r]);
        float* mg = v.cmarg + (size_t)ci * 36;
        float sum = 0.f;
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n2; ++j) { float q = P[i * 6 + j] * b1[i] * b2[j]; mg[i * 6 + j] = q; sum += q; }
        float is = 1.f / sum;
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n2; ++j) mg[i * 6 + j] *= is;
    }
    __syncthreads();
    return it;
}

// Bethe free energy (calculate_energy_from_marginals, :915-927).
__device__ float rot_free_energy(Ctx& c, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    float en = 0.f;
    for (int x = c.tid; x < R.n_node; x += NT) {
        int nr = v.nrot[x];
        float e = v.noff[x];
        for (int r = 0; r < nr; ++r) {
            float b = v.ncur[x * 6 + r];
            e += b * logf((1e-10f + b) / (1e-10f + v.nprob[x * 6 + r]));
        }
        en += e;
    }
    int total = R.n_node * v.maxc;
    for (int ci = c.tid; ci < total; ci += NT) {
        int a = ci / v.maxc, k = ci % v.maxc;
        if (k >= v.ccnt[a] || !v.cpres[ci]) continue;
        int b = v.cnbr[ci];
        const float* P = v.cP + (size_t)ci * 36;
        if (v.nrot[a] == 1 && v.nrot[b] == 1) { en += -logf(P[0]); continue; }
        int f, s;
        if (!rot_bp_edge(v, ci, f, s)) continue;
        int n1 = v.nrot[f], n2 = v.nrot[s];
        const float* mg = v.cmarg + (size_t)ci * 36;
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n2; ++j) {
                float p = mg[i * 6 + j];
                en += p * logf((1e-10f + p) / (1e-10f + P[i * 6 + j] * v.ncur[f * 6 + i] * v.ncur[s * 6 + j]));
            }
    }
    return block_sum(c, en);
}

// Edge sensitivities of bead pairs (src/rotamer.cpp:1057-1066) and the 1-body
// sensitivities (:1074-1085).
__device__ void rot_sens(Ctx& c, const PairGraph& G, const RotamerDev& R, const int* node_of_bead) {
    RotView v = rot_view(c, R);
    const int* cnt = c.I + G.cnt;
    const int* nbr = c.I + G.nbr;
    const int* act = c.I + G.act;
    float* es = c.A + G.esens;
    const int* loc1 = c.CI + G.loc1;
    int wave = c.tid >> 6, lane = c.tid & 63;
    for (int i = wave; i < G.n1; i += NWAVE) {
        int n = cnt[i];
        int a = node_of_bead[i];
        for (int k = lane; k < n; k += 64) {
            int slot = i * G.maxn + k;
            if (!act[slot]) { es[slot] = 0.f; continue; }
            int j = nbr[slot];
            int b = node_of_bead[j];
            int kk = lower_bound_i(v.cnbr + a * v.maxc, v.ccnt[a], b);
            if (kk >= v.ccnt[a] || v.cnbr[a * v.maxc + kk] != b) { es[slot] = 0.f; continue; }  // list overflow
            int ci = a * v.maxc + kk;
            bool sw = rot_swapped(v, a, b);
            int f = sw ? b : a, s = sw ? a : b;
            int rf = sw ? v.brot[j] : v.brot[i], rs = sw ? v.brot[i] : v.brot[j];
            float sv;
            if (v.nrot[f] == 1 && v.nrot[s] == 1) sv = 1.f;
            else if (v.nrot[f] == 1) sv = v.ncur[s * 6 + rs];
            else sv = v.cmarg[(size_t)ci * 36 + rf * 6 + rs];
            es[slot] = sv;
        }
    }
    for (int b = c.tid; b < G.n1; b += NT) {
        float m = v.ncur[node_of_bead[b] * 6 + v.brot[b]];
        for (int q = 0; q < R.n_prob; ++q) c.A[R.prob_sens[q] + loc1[b] * R.prob_w[q]] += m;
    }
    __syncthreads();
}

}  // namespace upside
