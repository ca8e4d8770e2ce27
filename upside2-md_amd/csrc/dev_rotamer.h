// Rotamer side-chain free energy by loopy belief propagation on the device
// (src/rotamer.cpp:628-1169).  Graph work is spread over the system's team of
// workgroups; the iterative solve runs on one workgroup when its state fits
// that workgroup's LDS, else on the whole team.
//
// Residue nodes carry 1, 3 or 6 rotamer states, one bead per state.  Residue
// pairs that may hold a bead pair of the rotamer pair graph's cache are
// "candidates" (kept per cache rebuild, in (a,b) order).  In an evaluation a candidate is
// present when one of its bead pairs is in the refined edge set, which is the
// reference's EdgeLocator edge set (src/rotamer.cpp:134-206,383-396):
//   * 1-rotamer x 1-rotamer pairs contribute -log P to the free energy;
//   * 1-rotamer x n pairs are folded into the n-rotamer node's probabilities
//     (move_edge_prob_to_node2, :398-405) -- here as the bead sum of their
//     edge energies, exp(-sum) being the product of the edge probabilities;
//   * 3/6 x 3/6 pairs become the BP edges, compacted per evaluation into
//     dense structure-of-arrays (matrix P, messages, marginals) so that one
//     thread updates one edge with coalesced loads.
// Messages are updated synchronously (Jacobi) per edge, node beliefs are
// products over incident edges renormalised after every factor, with the
// reference's schedule: an undamped first pass whose node beliefs are swapped
// out, then damped passes in chunks of `iteration_chunk_size`, stopping when
// the one-sided max change is below tol (src/rotamer.cpp:1089-1162).
#pragma once
#include "dev_pairs.h"

namespace upside {

struct RotView {
    gp<const int> nrot, bead0, brot, bnode;
    gp<int> coff, ca, cb, scand, ioff, iinc, eidx, bpc, bpf, bps, ninc, binc, stats;
    gp<float> P, msg, marg, e1x, nprob, ncur, nold, noff, feat, nprev;
    int cap;
};

__device__ RotView rot_view(Ctx& c, const RotamerDev& R) {
    RotView v;
    v.nrot = gmem(c.CI + R.node_nrot); v.bead0 = gmem(c.CI + R.node_bead0); v.brot = gmem(c.CI + R.bead_rot);
    v.bnode = gmem(c.CI + R.bead_node);
    v.coff = gmem(c.I + R.coff); v.ca = gmem(c.I + R.ca); v.cb = gmem(c.I + R.cb); v.scand = gmem(c.I + R.slot_cand);
    v.ioff = gmem(c.I + R.ioff); v.iinc = gmem(c.I + R.iinc); v.eidx = gmem(c.I + R.eidx);
    v.bpc = gmem(c.I + R.bpc); v.bpf = gmem(c.I + R.bpf); v.bps = gmem(c.I + R.bps); v.ninc = gmem(c.I + R.ninc);
    v.binc = gmem(c.I + R.binc); v.stats = gmem(c.I + R.stats);
    v.P = gmem(c.A + R.P); v.msg = gmem(c.A + R.msg); v.marg = gmem(c.A + R.marg); v.e1x = gmem(c.A + R.e1x);
    v.nprob = gmem(c.A + R.nprob); v.ncur = gmem(c.A + R.ncur); v.nold = gmem(c.A + R.nold);
    v.noff = gmem(c.A + R.noff);
    v.feat = gmem(c.A + R.feat);
    v.nprev = gmem(c.A + R.nprev);
    v.cap = R.ccap;
    return v;
}

// Residue-pair candidate lists from the node spheres cen (LDS or memory):
// counts, candidates (a < b, by a then b) and node incidences, each node
// handled by a 16-lane group testing 16 partners at a time (ballot
// compaction keeps partner order).  Returns false on capacity overflow.
template <class P>
__device__ bool rot_candidate_lists(Ctx& c, const PairGraph& G, const RotamerDev& R, const RotView& v, P cen) {
    const int n = R.n_node;
    const float rc = G.cache_cut;
    auto near = [&](int a, int b) {
        float dx = cen[a * 4] - cen[b * 4], dy = cen[a * 4 + 1] - cen[b * 4 + 1], dz = cen[a * 4 + 2] - cen[b * 4 + 2];
        float lim = (rc + cen[a * 4 + 3] + cen[b * 4 + 3]) * 1.0001f + 1e-3f;
        return dx * dx + dy * dy + dz * dz < lim * lim;
    };
    const int lane = c.tid & 63;
    const unsigned long long gmask = ((1ull << GW) - 1ull) << (lane & ~(GW - 1));
    const unsigned long long below = (1ull << lane) - 1ull;
    const int grp = c.gtid / GW, gl = c.gtid % GW, NG = c.gsize / GW;
    for (int base = 0; base < n; base += NG) {
        const int a = base + grp;
        const int lim = a < n ? n : 0;
        int up = 0, lo = 0;
        for (int b0 = 0; b0 < lim; b0 += GW) {
            const int b = b0 + gl;
            const bool nb = b < n && b != a && near(a, b);
            bool keep = nb && b > a;
            up += __popcll(__ballot(keep) & gmask);
            keep = nb && b < a;
            lo += __popcll(__ballot(keep) & gmask);
        }
        if (a < n && gl == 0) {
            v.coff[a] = up;
            v.ioff[a] = up + lo;
        }
    }
    if (c.gtid == 0) { v.coff[n] = 0; v.ioff[n] = 0; }
    team_sync(c);
    int C = team_exscan(c, v.coff, n + 1);
    int I = team_exscan(c, v.ioff, n + 1);
    if (C > R.ccap || I > 2 * R.ccap) {
        if (c.gtid == 0) { c.I[G.state + 1] = 1; c.I[G.state + 0] = 0; }
        team_sync(c);
        return false;
    }
    for (int base = 0; base < n; base += NG) {
        const int a = base + grp;
        const int lim = a < n ? n : 0;
        int k = a < n ? v.coff[a] : 0;
        for (int b0 = a + 1; b0 < lim; b0 += GW) {
            const int b = b0 + gl;
            bool keep = b < n && near(a, b);
            const unsigned long long m = __ballot(keep) & gmask;
            if (keep) {
                const int pos = k + __popcll(m & below);
                v.ca[pos] = a;
                v.cb[pos] = b;
            }
            k += __popcll(m);
        }
    }
    team_sync(c);
    for (int base = 0; base < n; base += NG) {
        const int x = base + grp;
        const int lim = x < n ? x : 0;
        int t = x < n ? v.ioff[x] : 0;
        for (int a0 = 0; a0 < lim; a0 += GW) {  // partners a < x, in increasing a
            const int a = a0 + gl;
            bool keep = a < x && near(a, x);
            int code = 0;
            if (keep) {
                int c0 = v.coff[a];
                code = (c0 + lower_bound_i(v.cb + c0, v.coff[a + 1] - c0, x)) * 2 + 1;
            }
            const unsigned long long m = __ballot(keep) & gmask;
            if (keep) v.iinc[t + __popcll(m & below)] = code;
            t += __popcll(m);
        }
        if (x < n)  // then its own candidates
            for (int cc = v.coff[x] + gl; cc < v.coff[x + 1]; cc += GW) v.iinc[t + cc - v.coff[x]] = cc * 2;
    }
    return true;
}

// Residue-pair candidates and node incidences, rebuilt with the bead rows.
// Candidates are the residue pairs whose bead spheres (centre = bead mean,
// radius = farthest bead) come within the cache radius of each other: a
// superset of the pairs with a bead pair in the cache, so every slot finds its
// candidate.  Candidate c = (ca[c], cb[c]), ca < cb, sorted by (a, b); the
// incidence list of node x names (candidate, 1) for partners a < x first, in
// increasing a, then (candidate, 0) for its own candidates.  Returns false on
// capacity overflow.
__device__ bool rot_rebuild(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    gp<const float> pack = gmem((const float*)c.A + G.pos_pack);
    gp<float> cen = gmem(c.A + R.ncen);
    gp<const int> row = gmem((const int*)c.I + G.row);
    gp<const int> nbr = gmem((const int*)c.I + G.nbr);
    const int n = R.n_node;
    for (int a = c.gtid; a < n; a += c.gsize) {
        int b0 = v.bead0[a], b1 = v.bead0[a + 1];
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int b = b0; b < b1; ++b) { sx += pack[b * 3]; sy += pack[b * 3 + 1]; sz += pack[b * 3 + 2]; }
        float inv = 1.f / (float)(b1 - b0);
        sx *= inv; sy *= inv; sz *= inv;
        float r2 = 0.f;
        for (int b = b0; b < b1; ++b) {
            float dx = pack[b * 3] - sx, dy = pack[b * 3 + 1] - sy, dz = pack[b * 3 + 2] - sz;
            r2 = fmaxf(r2, dx * dx + dy * dy + dz * dz);
        }
        cen[a * 4] = sx; cen[a * 4 + 1] = sy; cen[a * 4 + 2] = sz; cen[a * 4 + 3] = sqrtf(r2);
    }
    team_sync(c);
    // every workgroup reads the spheres from its own LDS copy when they fit
    bool ok;
    if (4 * n <= LDS_FLOATS) {
        lp<float> L = lmem(c.lds);
        for (int k = c.tid; k < 4 * n; k += NT) L[k] = cen[k];
        __syncthreads();
        ok = rot_candidate_lists(c, G, R, v, (lp<const float>)L);
    } else {
        ok = rot_candidate_lists(c, G, R, v, (gp<const float>)cen);
    }
    if (!ok) return false;
    // candidate of every bead slot
    const int E = c.I[G.state + 3];
    int bad = 0;
    for (int s = c.gtid; s < E; s += c.gsize) {
        int a = v.bnode[row[s]], b = v.bnode[nbr[s]];
        int c0 = v.coff[a], m = v.coff[a + 1] - c0;
        int k = lower_bound_i(v.cb + c0, m, b);
        if (k >= m || v.cb[c0 + k] != b) { bad = 1; k = 0; }
        v.scand[s] = c0 + k;
    }
    int anybad = team_or(c, bad);
    if (anybad) {
        if (c.gtid == 0) { c.I[G.state + 1] = 2; c.I[G.state + 0] = 0; }
        team_sync(c);
        return false;
    }
    if (c.gtid == 0) v.stats[5] = v.coff[n];  // candidate count
    team_sync(c);
    return true;
}

// Node probabilities, BP-edge compaction and edge matrices for this evaluation.
__device__ void rot_fill(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int* off = c.I + G.off;
    const int* nbr = c.I + G.nbr;
    const int* row = c.I + G.row;
    const int* toff = c.I + G.toff;
    const int* tslot = c.I + G.tslot;
    const int* act = c.I + G.act;
    const float* val = c.A + G.val;
    const int* loc1 = c.CI + G.loc1;
    const int E = c.I[G.state + 3];
    const int C = v.coff[R.n_node];
    // energies of each bead's edges to 1-rotamer residues (folded into the node)
    {
        int grp = c.gtid / GW, gl = c.gtid % GW;
        for (int base = 0; base < G.n1; base += c.gsize / GW) {
            int b = base + grp;
            bool ok = b < G.n1;
            float acc = 0.f;
            if (ok && v.nrot[v.bnode[b]] > 1) {
                for (int s = off[b] + gl; s < off[b + 1]; s += GW)
                    if (act[s] && v.nrot[v.bnode[nbr[s]]] == 1) acc += val[s];
                for (int t = toff[b] + gl; t < toff[b + 1]; t += GW) {
                    int s = tslot[t];
                    if (act[s] && v.nrot[v.bnode[row[s]]] == 1) acc += val[s];
                }
            }
            acc = group_sum<GW>(acc);
            if (ok && gl == 0) v.e1x[b] = acc;
        }
    }
    for (int k = c.gtid; k <= C; k += c.gsize) v.eidx[k] = 0;
    team_sync(c);
    // node energies -> probabilities (fill_holders, convert_energy_to_prob, :239-256,867-880)
    for (int n = c.gtid; n < R.n_node; n += c.gsize) {
        int b0 = v.bead0[n], nr = v.nrot[n];
        float e[6];
        float mn = 3.0e38f;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            float s = 0.f;
            if (k < nr) {
                for (int q = 0; q < R.n_prob; ++q) s += c.A[R.prob_out[q] + loc1[b0 + k] * R.prob_w[q]];
                mn = fminf(mn, s);
            }
            e[k] = s;
        }
        v.noff[n] = mn;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            v.nprob[n * 6 + k] = k < nr ? expf(mn - e[k] - v.e1x[b0 + k]) : 0.f;
    }
    // presence of BP edges (both residues with 3 or 6 rotamers)
    for (int s = c.gtid; s < E; s += c.gsize) {
        if (!act[s]) continue;
        if (v.nrot[v.bnode[row[s]]] > 1 && v.nrot[v.bnode[nbr[s]]] > 1) v.eidx[v.scand[s]] = 1;
    }
    team_sync(c);
    int nbp = team_exscan(c, v.eidx, C + 1);
    for (int cc = c.gtid; cc < C; cc += c.gsize) {
        int e = v.eidx[cc];
        if (v.eidx[cc + 1] == e) continue;
        int a = v.ca[cc], b = v.cb[cc];
        // rows of P belong to the residue with fewer rotamers (:898), the lower one on ties
        bool sw = v.nrot[a] > v.nrot[b];
        int f = sw ? b : a, s2 = sw ? a : b;
        v.bpc[e] = f | (s2 << 16); v.bpf[e] = f; v.bps[e] = s2;
        int n1 = v.nrot[f], n2 = v.nrot[s2];
#pragma unroll
        for (int q = 0; q < 36; ++q) v.P[e * 36 + q] = ((q / 6) < n1 && (q % 6) < n2) ? 1.f : 0.f;
    }
    if (c.gtid == 0) v.stats[4] = nbp;
    team_sync(c);
    // P[r1][r2] = exp(-E) of the bead pair (one bead per rotamer state)
    for (int s = c.gtid; s < E; s += c.gsize) {
        if (!act[s]) continue;
        int i = row[s], j = nbr[s];
        int a = v.bnode[i], b = v.bnode[j];
        if (v.nrot[a] == 1 || v.nrot[b] == 1) continue;
        int e = v.eidx[v.scand[s]];
        bool sw = v.nrot[a] > v.nrot[b];
        int ra = v.brot[i], rb = v.brot[j];
        int q = sw ? rb * 6 + ra : ra * 6 + rb;
        v.P[e * 36 + q] = expf(-val[s]);
    }
    // compact BP incidences of every node, in incidence order: a 16-lane
    // group per node tests 16 incidences at a time and keeps the BP ones
    // through a ballot
    {
        const int lane = c.tid & 63;
        const unsigned long long gmask = ((1ull << GW) - 1ull) << (lane & ~(GW - 1));
        const unsigned long long below = (1ull << lane) - 1ull;
        const int grp = c.gtid / GW, gl = c.gtid % GW, NG = c.gsize / GW;
        for (int base = 0; base < R.n_node; base += NG) {
            const int x = base + grp;
            const bool ok = x < R.n_node && v.nrot[x] > 1;
            const int t0 = ok ? v.ioff[x] : 0, t1 = ok ? v.ioff[x + 1] : 0;
            int k = t0;
            for (int t = t0; t < t1; t += GW) {  // same trip count for the group's lanes
                const int tt = t + gl;
                bool keep = false;
                int code = 0;
                if (tt < t1) {
                    int cc = v.iinc[tt] >> 1;
                    int e = v.eidx[cc];
                    keep = v.eidx[cc + 1] != e;
                    code = e * 2 + (keep && v.bpf[e] == x ? 0 : 1);
                }
                unsigned long long m = __ballot(keep) & gmask;
                if (keep) v.binc[k + __popcll(m & below)] = code;
                k += __popcll(m);
            }
            if (x < R.n_node && gl == 0) v.ninc[x] = k - t0;
        }
    }
    team_sync(c);
}

// Belief-propagation state of one solve.  Messages are updated in place (an
// edge's new messages depend only on its own old messages and on node
// beliefs), beliefs and incidences in LDS when they fit (F, I, J = LDS
// pointers), else in memory.
template <class F, class I, class J>
struct BpMem {
    // With the messages in memory (the team solve, and a workgroup solve too
    // large for LDS) every edge's 12 message components and 36 matrix entries
    // are contiguous (array of structures): a node gathers its incidences'
    // edges from scattered places, and so reads 1-2 cache lines per edge
    // instead of 48.  In LDS the edges are swept in order by consecutive
    // threads, and a structure of arrays (stride ms, ps) keeps those sweeps
    // conflict-free and coalesced.
    static constexpr bool AOS = !__is_same(F, lp<float>);
    F msg;             // component (side*6 + r) of edge e at msg[mi(side*6 + r, e)]
    int ms;
    gp<const float> P; // edge matrices (memory), entry (i, j) of edge e at P[pi(i*6 + j, e)]
    int ps;
    __device__ __forceinline__ int mi(int r, int e) const { return AOS ? e * 12 + r : r * ms + e; }
    __device__ __forceinline__ int pi(int q, int e) const { return AOS ? e * 36 + q : q * ps + e; }
    __device__ __forceinline__ int cs() const { return AOS ? 1 : ms; }  // stride of a message's components
    gp<const float> prob;  // node probabilities [n*6]
    I ends;            // edge endpoints: first | second << 16 (rows, columns of P)
    I ioff, ninc;      // node incidences inc[ioff[x] + t], t < ninc[x]
    J inc;             // incidence: edge*2 + side
    gp<const int> nrot;
    int c1, c2;        // edges [0,c1) are 3x3, [c1,c2) 3x6, the rest 6x6 (unsorted: c1 = c2 = 0)
    gp<const int> pos; // position of BP edge e in this order (null: the order of the BP edges)
};

// Who runs a BP pass: one workgroup (TEAM = false: threads c.tid of NT,
// workgroup barriers) or the system's whole team (TEAM = true: threads c.gtid
// of c.gsize, team barriers; state in memory).  The arithmetic of every edge
// and node is the same either way, so both give identical results.
template <bool TEAM> __device__ __forceinline__ int bp_t(const Ctx& c) { return TEAM ? c.gtid : c.tid; }
template <bool TEAM> __device__ __forceinline__ int bp_nt(const Ctx& c) { return TEAM ? c.gsize : NT; }
template <bool TEAM> __device__ __forceinline__ void bp_sync(Ctx& c) {
    if (TEAM) team_sync(c);
    else __syncthreads();
}
template <bool TEAM> __device__ __forceinline__ float bp_max(Ctx& c, float v) {
    return TEAM ? team_max(c, v) : block_max(c, v);
}

// New messages of edges e[0..W) (W = 1 or 2) of one class A x B (the first
// node has A <= B states; padded states of an edge never enter: their matrix
// entries and beliefs are zero) from old node beliefs `nold_` and the edges'
// old messages; each new message is scaled to unit L1 norm
// (src/rotamer.cpp:515-568).  Every load of the W edges is issued before any
// of their stores, so the loads are in flight together.
template <int W, int A, int B, class F, class I, class J>
__device__ __forceinline__ void rot_edge_w(const BpMem<F, I, J>& m, const int* e, F nold_) {
    float p[W][A * B], v1[W][A], v2[W][B];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int ends = m.ends[e[w]];
        const int f = ends & 0xFFFF, s = ends >> 16;
#pragma unroll
        for (int i = 0; i < A; ++i)
#pragma unroll
            for (int j = 0; j < B; ++j) p[w][i * B + j] = m.P[m.pi(i * 6 + j, e[w])];
        // belief / old message by the hardware reciprocal (1 ulp) rather
        // than a correctly rounded division: these dominate the pass
#pragma unroll
        for (int r = 0; r < A; ++r)
            v1[w][r] = nold_[f * 6 + r] * __builtin_amdgcn_rcpf(1e-10f + m.msg[m.mi(r, e[w])]);
#pragma unroll
        for (int r = 0; r < B; ++r)
            v2[w][r] = nold_[s * 6 + r] * __builtin_amdgcn_rcpf(1e-10f + m.msg[m.mi(6 + r, e[w])]);
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        float m1[A], m2[B];
#pragma unroll
        for (int i = 0; i < A; ++i) m1[i] = 0.f;
#pragma unroll
        for (int j = 0; j < B; ++j) m2[j] = 0.f;
#pragma unroll
        for (int i = 0; i < A; ++i)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                m1[i] = __fmaf_rn(p[w][i * B + j], v2[w][j], m1[i]);
                m2[j] = __fmaf_rn(p[w][i * B + j], v1[w][i], m2[j]);
            }
        float s1 = 0.f, s2 = 0.f;
#pragma unroll
        for (int i = 0; i < A; ++i) s1 = __fadd_rn(s1, m1[i]);
#pragma unroll
        for (int j = 0; j < B; ++j) s2 = __fadd_rn(s2, m2[j]);
        float i1 = 1.f / s1, i2 = 1.f / s2;
#pragma unroll
        for (int i = 0; i < A; ++i) m.msg[m.mi(i, e[w])] = m1[i] * i1;
#pragma unroll
        for (int j = 0; j < B; ++j) m.msg[m.mi(6 + j, e[w])] = m2[j] * i2;
    }
}

template <int W, class F, class I, class J>
__device__ __forceinline__ void rot_edge_cls(const BpMem<F, I, J>& m, int cls, const int* e, F nold_) {
    if (cls == 0) rot_edge_w<W, 3, 3>(m, e, nold_);
    else if (cls == 1) rot_edge_w<W, 3, 6>(m, e, nold_);
    else rot_edge_w<W, 6, 6>(m, e, nold_);
}

// One synchronous message pass over all BP edges.  A thread takes edges e and
// e+nt together when they are of one class (edges sorted by class keep most
// waves on one code path).
template <bool TEAM, class F, class I, class J>
__device__ void rot_edge_pass(Ctx& c, const BpMem<F, I, J>& m, int nbp, F nold_) {
    auto cls = [&](int k) { return k < m.c1 ? 0 : (k < m.c2 ? 1 : 2); };
    const int nt = bp_nt<TEAM>(c);
    int e = bp_t<TEAM>(c);
    for (; e + nt < nbp; e += 2 * nt) {
        const int pair[2] = {e, e + nt};
        const int ca = cls(e), cb = cls(e + NT);
        if (ca == cb) {
            rot_edge_cls<2>(m, ca, pair, nold_);
        } else {
            rot_edge_cls<1>(m, ca, pair, nold_);
            rot_edge_cls<1>(m, cb, pair + 1, nold_);
        }
    }
    if (e < nbp) rot_edge_cls<1>(m, cls(e), &e, nold_);
    bp_sync<TEAM>(c);
}

// Node pass: cur = prob * prod(new messages), renormalised after every factor;
// then cur = (1-d) cur/max(cur) + d*old (standardize_belief_update, :258-273).
// With swap_first, the product goes to `nold_` and cur restarts from prob
// (the first, undamped pass of solve_for_marginals).  Padding states carry
// zeros throughout.
//   On one workgroup a thread takes a node (incidences and messages in LDS
// are cheap to walk).  On the team, with the state in memory, a group of
// NODE_LANES lanes takes a node: each lane loads one incidence and its
// message, and every lane of the group then applies the factors in incidence
// order from the group's registers, so the arithmetic (and the result) is the
// same as a single thread's while the dependent memory trips are paid once
// per NODE_LANES incidences.
constexpr int NODE_LANES = 16;

// The roundings are pinned (no contraction of a product into a later sum) so
// that every code path computing a node product gives the same bits.
template <class F>
__device__ __forceinline__ void bp_factor(float* b, const float* a) {
    float sum = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) { b[r] = __fmul_rn(b[r], a[r]); sum = __fadd_rn(sum, b[r]); }
    float inv = __builtin_amdgcn_rcpf(sum);  // renormalisation only: 1 ulp is enough
#pragma unroll
    for (int r = 0; r < 6; ++r) b[r] = __fmul_rn(b[r], inv);
}

template <class F>
__device__ __forceinline__ void bp_node_store(int x, const float* pr, const float* b, F ncur_, F nold_,
                                              float damping, bool swap_first) {
    if (swap_first) {
        float mx = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) mx = fmaxf(mx, pr[r]);
        float im = 1.f / mx;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            nold_[x * 6 + r] = b[r];
            ncur_[x * 6 + r] = pr[r] * im;
        }
    } else {
        float mx = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) mx = fmaxf(mx, b[r]);
        float im = 1.f / mx;
        const float w = __fmul_rn(1.f - damping, im);
#pragma unroll
        for (int r = 0; r < 6; ++r) ncur_[x * 6 + r] = __fmaf_rn(damping, nold_[x * 6 + r], __fmul_rn(w, b[r]));
    }
}

template <bool TEAM, class F, class I, class J>
__device__ void rot_node_pass(Ctx& c, const BpMem<F, I, J>& m, int n_node, F ncur_, F nold_, float damping,
                              bool swap_first) {
    if (TEAM) {
        const int t = bp_t<TEAM>(c), gl = t % NODE_LANES;
        for (int x = t / NODE_LANES; x < n_node; x += bp_nt<TEAM>(c) / NODE_LANES) {  // uniform in a group
            if (m.nrot[x] == 1) continue;
            float pr[6], b[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) { pr[r] = m.prob[x * 6 + r]; b[r] = pr[r]; }
            const int t0 = m.ioff[x], n = m.ninc[x];
            for (int tb = 0; tb < n; tb += NODE_LANES) {
                float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (tb + gl < n) {
                    const int code = m.inc[t0 + tb + gl];
                    const int b0 = m.mi((code & 1) * 6, code >> 1);
#pragma unroll
                    for (int r = 0; r < 6; ++r) a[r] = m.msg[b0 + r * m.cs()];
                }
                const int cnt = min(NODE_LANES, n - tb);
                for (int q = 0; q < cnt; ++q) {
                    float f[6];
#pragma unroll
                    for (int r = 0; r < 6; ++r) f[r] = __shfl(a[r], q, NODE_LANES);
                    bp_factor<F>(b, f);
                }
            }
            if (gl == 0) bp_node_store(x, pr, b, ncur_, nold_, damping, swap_first);
        }
        bp_sync<TEAM>(c);
        return;
    }
    for (int x = bp_t<TEAM>(c); x < n_node; x += bp_nt<TEAM>(c)) {
        if (m.nrot[x] == 1) continue;
        float pr[6], b[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) { pr[r] = m.prob[x * 6 + r]; b[r] = pr[r]; }
        int t0 = m.ioff[x], n = m.ninc[x];
        // four incidences per step: their indices, then all their messages,
        // are loaded before the first factor is applied (the factors stay in
        // incidence order)
        int t = 0;
        for (; t + 3 < n; t += 4) {
            int base[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int code = m.inc[t0 + t + q];
                base[q] = m.mi((code & 1) * 6, code >> 1);
            }
            float a[4][6];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 6; ++r) a[q][r] = m.msg[base[q] + r * m.cs()];
#pragma unroll
            for (int q = 0; q < 4; ++q) bp_factor<F>(b, a[q]);
        }
        for (; t < n; ++t) {
            const int code = m.inc[t0 + t];
            const int b0 = m.mi((code & 1) * 6, code >> 1);
            float a0[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) a0[r] = m.msg[b0 + r * m.cs()];
            bp_factor<F>(b, a0);
        }
        bp_node_store(x, pr, b, ncur_, nold_, damping, swap_first);
    }
    bp_sync<TEAM>(c);
}

// One synchronous iteration of the team solve in a single phase: a group of
// NODE_LANES lanes takes a node, each lane one of its incidences, and computes
// the edge's new message into that node from the old messages (min) and the
// old belief of the edge's other node, in the summation order rot_edge_w
// uses, and stores it into mout; the group multiplies the messages into its
// node in incidence order as rot_node_pass does.  Each message is computed
// once, by the group of the node it goes to, and the message and node passes
// need one team barrier instead of two.  Results equal rot_edge_pass +
// rot_node_pass.
template <class F, class I, class J>
__device__ void rot_fused_pass(Ctx& c, const BpMem<F, I, J>& m, gp<const float> min_, gp<float> mout, int n_node,
                               F ncur_, F nold_, float damping, bool swap_first) {
    const int t = c.gtid, gl = t % NODE_LANES;
    for (int x = t / NODE_LANES; x < n_node; x += c.gsize / NODE_LANES) {  // uniform in a group
        if (m.nrot[x] == 1) continue;
        float pr[6], b[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) { pr[r] = m.prob[x * 6 + r]; b[r] = pr[r]; }
        const int t0 = m.ioff[x], n = m.ninc[x];
        for (int tb = 0; tb < n; tb += NODE_LANES) {
            float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (tb + gl < n) {
                const int code = m.inc[t0 + tb + gl];
                // the lane computes only the message into its own node: side 0
                // (the edge's first node f) m1 = P v2 from s, side 1 m2 = P^T v1
                // from f, each summed in the order rot_edge_w uses, and stores it
                const int e = code >> 1, side = code & 1;
                const int ends = m.ends[e];
                const int from = side ? (ends & 0xFFFF) : (ends >> 16);
                float p[36], v[6], mm[6];
#pragma unroll
                for (int q = 0; q < 36; ++q) p[q] = m.P[m.pi(q, e)];
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    v[r] = nold_[from * 6 + r] * __builtin_amdgcn_rcpf(1e-10f + min_[m.mi((side ? 0 : 6) + r, e)]);
                    mm[r] = 0.f;
                }
#pragma unroll
                for (int i = 0; i < 6; ++i)
#pragma unroll
                    for (int j = 0; j < 6; ++j)
                        mm[i] = __fmaf_rn(side ? p[j * 6 + i] : p[i * 6 + j], v[j], mm[i]);
                float sm = 0.f;
#pragma unroll
                for (int r = 0; r < 6; ++r) sm = __fadd_rn(sm, mm[r]);
                const float im = 1.f / sm;
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    a[r] = mm[r] * im;
                    mout[m.mi((side ? 6 : 0) + r, e)] = a[r];
                }
            }
            const int cnt = min(NODE_LANES, n - tb);
            for (int q = 0; q < cnt; ++q) {
                float fa[6];
#pragma unroll
                for (int r = 0; r < 6; ++r) fa[r] = __shfl(a[r], q, NODE_LANES);
                bp_factor<F>(b, fa);
            }
        }
        if (gl == 0) bp_node_store(x, pr, b, ncur_, nold_, damping, swap_first);
    }
    team_sync(c);
}

// ---- warm start (bp_mode 1 and 2; no counterpart in the reference) ----------
// Features of node x after the first pass (b0 = its undamped product of the
// first messages, pr = its probabilities, prev = its marginal at the end of
// the previous solve of this system, all zero before the first), MLP_IN floats:
//   [0,6)   log pr_r - max (clamped at -30), 0 off the node's states
//   [6,12)  log b0_r - max (clamped at -30): the first pass's log beliefs
//   [12,18) log prev_r - max (clamped at -30), 0 without a previous solve
//   [18,24) state mask
//   [24]    BP incidences / 8 (edge degree), [25] six states, [26] entropy
//   and [27] 1 - max of softmax([6,12)), [28] a previous solve exists, the
//   rest 0.
// Nodes with one state get a row of zeros.
template <bool TEAM, class F, class I, class J>
__device__ void bp_features(Ctx& c, const BpMem<F, I, J>& m, int n_node, F b0, gp<const float> prev,
                            gp<float> feat) {
    for (int x = bp_t<TEAM>(c); x < n_node; x += bp_nt<TEAM>(c)) {
        float f[MLP_IN];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) f[k] = 0.f;
        const int nr = m.nrot[x];
        if (nr > 1) {
            float mp = -3.0e38f, mb = -3.0e38f;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r < nr) {
                    f[r] = logf(fmaxf(m.prob[x * 6 + r], 1e-30f));
                    f[6 + r] = logf(fmaxf(b0[x * 6 + r], 1e-30f));
                    mp = fmaxf(mp, f[r]);
                    mb = fmaxf(mb, f[6 + r]);
                }
            float se = 0.f, sle = 0.f;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r < nr) {
                    f[r] = fmaxf(f[r] - mp, -30.f);
                    f[6 + r] = fmaxf(f[6 + r] - mb, -30.f);
                    f[18 + r] = 1.f;
                    const float e = expf(f[6 + r]);
                    se += e;
                    sle += e * f[6 + r];
                }
            f[24] = (float)m.ninc[x] * 0.125f;
            f[25] = nr == 6 ? 1.f : 0.f;
            f[26] = logf(se) - sle / se;   // entropy of softmax(log b0)
            f[27] = 1.f - 1.f / se;        // 1 - its largest probability (max log b0 is 0)
            float pv[6], mv = 0.f;
#pragma unroll
            for (int r = 0; r < 6; ++r) { pv[r] = prev[x * 6 + r]; mv = fmaxf(mv, pv[r]); }
            if (mv > 0.f) {
                const float lm = logf(mv);
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r < nr) f[12 + r] = fmaxf(logf(fmaxf(pv[r], 1e-30f)) - lm, -30.f);
                f[28] = 1.f;
            }
        }
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) feat[x * MLP_IN + k] = f[k];
    }
    bp_sync<TEAM>(c);
}

// The network over all node rows (16-row tiles, one wave each): the warm
// beliefs of node x are exp(z_r - max z), z_r = feat[6 + r] + out_r over its
// states (max-normalised, as the damped update leaves beliefs), into cur.
template <bool TEAM, class F, class I, class J>
__device__ void bp_warm_beliefs(Ctx& c, const BpMem<F, I, J>& m, const MlpDev& w, int n_node,
                                gp<const float> feat, F cur) {
    const int t = bp_t<TEAM>(c), nw = bp_nt<TEAM>(c) / 64;
    const int l = t & 63, g = l >> 4, cl = l & 15;
    for (int tile = t / 64; tile * 16 < n_node; tile += nw) {  // uniform over the wave
        f32x4 y;
        mlp_tile_infer(w, feat, n_node, tile * 16, y);
        float yo[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            yo[i] = __shfl(y[i], cl, 64);
            yo[4 + i] = __shfl(y[i], 16 + cl, 64);
        }
        const int x = tile * 16 + cl;
        if (g == 0 && x < n_node) {
            const int nr = m.nrot[x];
            if (nr > 1) {
                float z[6], mx = -3.0e38f;
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    z[r] = feat[x * MLP_IN + 6 + r] + yo[r];
                    if (r < nr) mx = fmaxf(mx, z[r]);
                }
#pragma unroll
                for (int r = 0; r < 6; ++r) cur[x * 6 + r] = r < nr ? expf(z[r] - mx) : 0.f;
            }
        }
    }
    bp_sync<TEAM>(c);
}

// Messages consistent with given node beliefs b: per edge, K updates of its
// two messages in registers with the beliefs held fixed (each new message
// from the other side's cavity belief b / message, scaled to unit L1 norm,
// as in rot_edge_w).  Every edge on the 6x6 path (padding is zero).
template <bool TEAM, class F, class I, class J>
__device__ void bp_local_messages(Ctx& c, const BpMem<F, I, J>& m, int nbp, F b, int K) {
    for (int e = bp_t<TEAM>(c); e < nbp; e += bp_nt<TEAM>(c)) {
        const int ends = m.ends[e];
        const int f = ends & 0xFFFF, s = ends >> 16;
        float p[36], m1[6], m2[6], bf[6], bs[6];
#pragma unroll
        for (int q = 0; q < 36; ++q) p[q] = m.P[m.pi(q, e)];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            m1[r] = m.msg[m.mi(r, e)];
            m2[r] = m.msg[m.mi(6 + r, e)];
            bf[r] = b[f * 6 + r];
            bs[r] = b[s * 6 + r];
        }
        for (int k = 0; k < K; ++k) {
            float v1[6], v2[6], n1[6], n2[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                v1[r] = bf[r] * __builtin_amdgcn_rcpf(1e-10f + m1[r]);
                v2[r] = bs[r] * __builtin_amdgcn_rcpf(1e-10f + m2[r]);
                n1[r] = 0.f;
                n2[r] = 0.f;
            }
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    n1[i] += p[i * 6 + j] * v2[j];
                    n2[j] += p[i * 6 + j] * v1[i];
                }
            float s1 = 0.f, s2 = 0.f;
#pragma unroll
            for (int r = 0; r < 6; ++r) { s1 += n1[r]; s2 += n2[r]; }
            const float i1 = 1.f / s1, i2 = 1.f / s2;
#pragma unroll
            for (int r = 0; r < 6; ++r) { m1[r] = n1[r] * i1; m2[r] = n2[r] * i2; }
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            m.msg[m.mi(r, e)] = m1[r];
            m.msg[m.mi(6 + r, e)] = m2[r];
        }
    }
    bp_sync<TEAM>(c);
}

// The plain start: messages 1, old beliefs = probabilities (:1110-1133).
template <bool TEAM, class F, class I, class J>
__device__ void bp_plain_start(Ctx& c, const BpMem<F, I, J>& m, int nbp, int n_node, F old) {
    for (int e = bp_t<TEAM>(c); e < nbp; e += bp_nt<TEAM>(c))
#pragma unroll
        for (int r = 0; r < 12; ++r) m.msg[m.mi(r, e)] = 1.f;
    for (int i = bp_t<TEAM>(c); i < 6 * n_node; i += bp_nt<TEAM>(c)) old[i] = m.prob[i];
    bp_sync<TEAM>(c);
}

// Iterations, node marginals and edge marginals of one solve whose state is
// set up in m (messages all 1, `old` = prob); returns the iteration count.
//   Schedules (RunArgs::bp_mode): 0 is the reference's; 1 stops instead
// when the largest two-sided change of a belief over a chunk is at most
// bp_residual; 2 replaces the beliefs after the first pass by the
// warm-start network's and derives consistent messages (bp_local_messages),
// stops by the residual rule, and when bp_budget iterations do not meet it
// restarts from the plain start as schedule 1 (the iteration count returned
// includes both).  Schedule 3 is the warm start of schedule 2 stopped by the
// reference's rule (the largest one-sided change at most R.tol, or at most
// bp_residual when that is positive), with the same budget and plain
// restart.  Schedules 1-3 also leave the features of the
// first pass in R.feat (training rows for the network).
template <bool TEAM, class F, class I, class J>
__device__ int bp_iterate(Ctx& c, const RotamerDev& R, const RotView& v, const BpMem<F, I, J>& m, int nbp, F cur,
                          F old) {
    const int n = R.n_node;
    const int cap = v.cap;
    const int t = bp_t<TEAM>(c), nt = bp_nt<TEAM>(c);
    const int mode = c.args->bp_mode;
    bool warm = mode == 2 || mode == 3;
    const bool resid = mode == 1 || mode == 2;  // residual rule; else the reference's
    int total = 0;
    // On the team the message and node passes of an iteration are fused
    // (rot_fused_pass) over two message buffers that alternate; mm.msg is the
    // buffer holding the current messages.
    BpMem<F, I, J> mm = m;
    F other{};
    if constexpr (TEAM) other = gmem(c.A + R.msg2);
    // The first (undamped) pass writes the product into `old`, which the
    // fused pass's message updates are still reading: it keeps two phases.
    auto iterate = [&](F cur_, F old_, float damping, bool swap_first) {
        unsigned long long t0 = ptick(c);
        if constexpr (TEAM) {
            if (!swap_first) {
                rot_fused_pass(c, mm, (gp<const float>)mm.msg, (gp<float>)other, n, cur_, old_, damping, false);
                F tt = mm.msg; mm.msg = other; other = tt;
                pacc(c, PROF_BP_NODE, t0);
                return;
            }
        }
        rot_edge_pass<TEAM>(c, mm, nbp, old_);
        pacc(c, PROF_BP_EDGE, t0);
        t0 = ptick(c);
        rot_node_pass<TEAM>(c, mm, n, cur_, old_, damping, swap_first);
        pacc(c, PROF_BP_NODE, t0);
    };
    for (;;) {
        iterate(cur, old, 0.f, true);
        if (mode > 0 && total == 0)  // old: the first pass's product
            bp_features<TEAM>(c, mm, n, old, (gp<const float>)v.nprev, v.feat);
        if (warm) {
            bp_warm_beliefs<TEAM>(c, mm, c.args->mlp, n, (gp<const float>)v.feat, cur);
            bp_local_messages<TEAM>(c, mm, nbp, cur, c.args->bp_klocal);
        }
        const float tol = resid || (mode == 3 && c.args->bp_residual > 0.f) ? c.args->bp_residual : R.tol;
        const int limit = warm ? min(c.args->bp_budget, R.max_iter) : R.max_iter;
        float dev = 1e10f;
        int it = 0;
        for (; dev > tol && it < limit; it += R.chunk) {
            for (int j = 0; j < R.chunk; ++j) {
                F tt = cur; cur = old; old = tt;          // swap node beliefs
                iterate(cur, old, R.damping, false);
            }
            float d = 0.f;
            for (int x = t; x < n; x += nt) {
                if (m.nrot[x] == 1) continue;
                if (resid) {
#pragma unroll
                    for (int r = 0; r < 6; ++r) d = fmaxf(d, fabsf(cur[x * 6 + r] - old[x * 6 + r]));
                } else {
#pragma unroll
                    for (int r = 0; r < 6; ++r) d = fmaxf(d, cur[x * 6 + r] - old[x * 6 + r]);
                }
            }
            dev = bp_max<TEAM>(c, d);
        }
        total += it;
        if (c.gtid == 0) {
            if (warm) v.stats[7] += 1;
            if (resid && dev <= tol) v.stats[9] += 1;
        }
        if (!warm || dev <= tol) break;
        // the warm start missed the residual within its budget: plain start
        if (c.gtid == 0) v.stats[8] += 1;
        warm = false;
        bp_plain_start<TEAM>(c, mm, nbp, n, old);
    }
    const int it = total;
    unsigned long long t0 = ptick(c);
    // node marginals (:283-290) into memory and into `old` (free now; cur and
    // old are always distinct buffers)
    gp<float> ncur = gmem(v.ncur);
    for (int x = t; x < n; x += nt) {
        int nr = m.nrot[x];
        float b[6];
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) { b[r] = cur[x * 6 + r]; s += b[r]; }
        float is = 1.f / s;
#pragma unroll
        for (int r = 0; r < 6; ++r) b[r] = nr == 1 ? (r == 0 ? 1.f : 0.f) : b[r] * is;
#pragma unroll
        for (int r = 0; r < 6; ++r) { ncur[x * 6 + r] = b[r]; old[x * 6 + r] = b[r]; v.nprev[x * 6 + r] = b[r]; }
    }
    bp_sync<TEAM>(c);
    // edge marginals (:425-449), written in the order of the BP edges
    gp<float> marg = gmem(v.marg);
    for (int e = t; e < nbp; e += nt) {
        const int k = m.pos ? m.pos[e] : e;
        const int ends = m.ends[k];
        const int f = ends & 0xFFFF, s = ends >> 16;
        float b1[6], b2[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            b1[r] = old[f * 6 + r] / (1e-10f + mm.msg[m.mi(r, k)]);
            b2[r] = old[s * 6 + r] / (1e-10f + mm.msg[m.mi(6 + r, k)]);
        }
        float p[36];
#pragma unroll
        for (int q = 0; q < 36; ++q) p[q] = m.P[m.pi(q, k)];
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) sum += p[i * 6 + j] * b1[i] * b2[j];
        float is = 1.f / sum;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) marg[(i * 6 + j) * cap + e] = p[i * 6 + j] * b1[i] * b2[j] * is;
    }
    pacc(c, PROF_BP_SETUP, t0);
    return it;
}

// LDS floats a solve in LDS needs: messages 12 per edge, packed endpoints 1
// per edge, incidences (16-bit, two per edge) 1 per edge, beliefs 12 per
// node, incidence offsets and counts 2 per node.  The message area doubles as
// scratch for sorting the edges by class while staging (5 per edge).
__device__ __forceinline__ int bp_lds_floats(int nbp, int n) { return 14 * nbp + 14 * n + 2; }

// Whether a solve of nbp BP edges over n nodes runs on one workgroup with
// its state in LDS (else on the whole team from memory, or, with
// bp_layout 2, on one workgroup from memory).
__device__ __forceinline__ bool bp_fits_lds(int nbp, int n, int cap) {
    return bp_lds_floats(nbp, n) <= LDS_FLOATS && 2 * nbp < 65536 && 36 * nbp <= 12 * cap;
}

// The solve proper, on one workgroup: a pass has a few thousand flops per
// node, too little to pay for team barriers, and a single workgroup can keep
// messages, beliefs and incidences in LDS (then every dependent load of the
// node products is an LDS read).  In LDS the edges are sorted by class
// (3x3, 3x6, 6x6; stable) and their matrices copied in that order into the
// memory message area, which the LDS solve does not use.  Returns the
// iteration count.
__device__ int rot_solve_wg(Ctx& c, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int nbp = v.stats[4];
    const int n = R.n_node;
    const int cap = v.cap;
    unsigned long long t0 = ptick(c);
    if (bp_fits_lds(nbp, n, cap)) {
        if (c.tid == 0) v.stats[6] += 1;  // solves held in LDS
        lp<float> L = lmem(c.lds);
        BpMem<lp<float>, lp<int>, lp<unsigned short>> m;
        m.nrot = v.nrot; m.prob = (gp<const float>)v.nprob;
        m.msg = L; m.ms = nbp; L += 12 * nbp;
        lp<float> cur = L; L += 6 * n;
        lp<float> old = L; L += 6 * n;
        lp<int> Ii = (lp<int>)L;
        lp<int> ends = Ii; Ii += nbp;
        lp<int> ioff = Ii; Ii += n + 1;
        lp<int> ninc = Ii; Ii += n;
        lp<unsigned short> inc = (lp<unsigned short>)Ii;  // 2*nbp entries
        // sort scratch in the message area: class indicators [3][nbp], stashed
        // endpoints [nbp], positions [nbp]
        lp<int> S = (lp<int>)m.msg;
        lp<int> stash = S + 3 * nbp;
        lp<int> posl = S + 4 * nbp;
        for (int k = c.tid; k < 6 * n; k += NT) old[k] = v.nprob[k];
        for (int x = c.tid; x < n; x += NT) { int k = v.ninc[x]; ninc[x] = k; ioff[x] = k; }
        if (c.tid == 0) ioff[n] = 0;
        for (int e = c.tid; e < nbp; e += NT) {
            const int en = v.bpc[e];
            const int nf = v.nrot[en & 0xFFFF], ns = v.nrot[en >> 16];
            const int cls = nf == 6 ? 2 : (ns == 6 ? 1 : 0);
            S[e] = cls == 0;
            S[nbp + e] = cls == 1;
            S[2 * nbp + e] = cls == 2;
            stash[e] = en;
        }
        __syncthreads();
        block_exscan(c, ioff, n + 1);
        const int n0 = block_exscan(c, S, nbp);
        const int n1 = block_exscan(c, S + nbp, nbp);
        block_exscan(c, S + 2 * nbp, nbp);
        m.c1 = n0;
        m.c2 = n0 + n1;
        gp<float> Ps = v.msg;  // matrices in class order, stride nbp
        gp<int> gpos = v.bpc;  // position of every BP edge (endpoints now staged)
        for (int e = c.tid; e < nbp; e += NT) {
            const int en = stash[e];
            const int nf = v.nrot[en & 0xFFFF], ns = v.nrot[en >> 16];
            const int cls = nf == 6 ? 2 : (ns == 6 ? 1 : 0);
            const int k = (cls == 0 ? 0 : cls == 1 ? m.c1 : m.c2) + S[cls * nbp + e];
            ends[k] = en;
            posl[e] = k;
            gpos[e] = k;
#pragma unroll 4
            for (int q = 0; q < 36; ++q) Ps[q * nbp + k] = v.P[e * 36 + q];
        }
        __syncthreads();
        for (int x = c.tid; x < n; x += NT) {
            int s0 = v.ioff[x], d0 = ioff[x], k = ninc[x];
            for (int t = 0; t < k; ++t) {
                int code = v.binc[s0 + t];
                inc[d0 + t] = (unsigned short)(posl[code >> 1] * 2 + (code & 1));
            }
        }
        __syncthreads();  // the scratch is overwritten by the messages next
        m.P = (gp<const float>)Ps; m.ps = nbp; m.pos = (gp<const int>)gpos;
        m.ends = ends; m.ioff = ioff; m.ninc = ninc; m.inc = inc;
        // initial messages = 1, old beliefs = prob (:1110-1133)
        for (int e = c.tid; e < nbp; e += NT)
#pragma unroll
            for (int r = 0; r < 12; ++r) m.msg[m.mi(r, e)] = 1.f;
        __syncthreads();
        pacc(c, PROF_BP_SETUP, t0);
        return bp_iterate<false>(c, R, v, m, nbp, cur, old);
    }
    BpMem<gp<float>, gp<int>, gp<int>> m;
    m.P = (gp<const float>)v.P; m.ps = 0; m.nrot = v.nrot; m.prob = (gp<const float>)v.nprob;
    m.msg = v.msg; m.ms = 0;
    m.ends = v.bpc; m.ioff = v.ioff; m.ninc = v.ninc; m.inc = v.binc;
    m.c1 = m.c2 = 0;
    m.pos = nullptr;
    for (int i = c.tid; i < n * 6; i += NT) v.nold[i] = v.nprob[i];
    for (int e = c.tid; e < nbp; e += NT)
#pragma unroll
        for (int r = 0; r < 12; ++r) m.msg[m.mi(r, e)] = 1.f;
    __syncthreads();
    pacc(c, PROF_BP_SETUP, t0);
    return bp_iterate<false>(c, R, v, m, nbp, v.ncur, v.nold);
}

// The solve on the whole team with its state in memory (matrices, messages,
// beliefs, incidences): every pass is spread over all the team's threads and
// separated by team barriers.  Every thread of the team returns the
// iteration count.
__device__ int rot_solve_team(Ctx& c, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int nbp = v.stats[4];
    const int n = R.n_node;
    unsigned long long t0 = ptick(c);
    BpMem<gp<float>, gp<int>, gp<int>> m;
    m.P = (gp<const float>)v.P; m.ps = 0; m.nrot = v.nrot; m.prob = (gp<const float>)v.nprob;
    m.msg = v.msg; m.ms = 0;
    m.ends = v.bpc; m.ioff = v.ioff; m.ninc = v.ninc; m.inc = v.binc;
    m.c1 = m.c2 = 0;
    m.pos = nullptr;
    for (int i = c.gtid; i < n * 6; i += c.gsize) v.nold[i] = v.nprob[i];
    for (int e = c.gtid; e < nbp; e += c.gsize)
#pragma unroll
        for (int r = 0; r < 12; ++r) m.msg[m.mi(r, e)] = 1.f;
    team_sync(c);
    pacc(c, PROF_BP_SETUP, t0);
    return bp_iterate<true>(c, R, v, m, nbp, v.ncur, v.nold);
}

// Where a solve runs: true = rank 0 alone (rot_solve_wg), false = several
// workgroups (rot_solve_team: the whole team, or its first half beside the
// independent ops).  bp_layout 0: one workgroup when the state fits LDS or
// the team has one workgroup, else several; 1: always the whole team; 2:
// always one workgroup; 3: always several.  Uniform over the team (nbp is
// read after a team barrier).
__device__ __forceinline__ bool bp_on_one_wg(const Ctx& c, const RotamerDev& R, int nbp, int bp_layout) {
    if (c.G == 1 || bp_layout == 2) return true;
    if (bp_layout == 1 || bp_layout == 3) return false;
    return bp_fits_lds(nbp, R.n_node, R.ccap);
}

// Solve by rank 0 of the team while the others wait at the barrier, or by
// the whole team (see bp_on_one_wg).
__device__ int rot_solve(Ctx& c, const RotamerDev& R, int bp_layout) {
    const int nbp = c.I[R.stats + 4];
    if (!bp_on_one_wg(c, R, nbp, bp_layout)) return rot_solve_team(c, R);
    int it = 0;
    if (c.rank == 0) it = rot_solve_wg(c, R);
    team_sync(c);
    return it;
}

// Bethe free energy (calculate_energy_from_marginals, :915-927).
__device__ float rot_free_energy(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int cap = v.cap;
    float en = 0.f;
    for (int x = c.gtid; x < R.n_node; x += c.gsize) {
        float e = v.noff[x];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            float b = v.ncur[x * 6 + r];
            e += b * logf((1e-10f + b) / (1e-10f + v.nprob[x * 6 + r]));
        }
        en += e;
    }
    // 1-rotamer pairs: -log P with P the probability of their single bead pair
    const int* row = c.I + G.row;
    const int* nbr = c.I + G.nbr;
    const int* act = c.I + G.act;
    const float* val = c.A + G.val;
    const int E = c.I[G.state + 3];
    for (int s = c.gtid; s < E; s += c.gsize)
        if (act[s] && v.nrot[v.bnode[row[s]]] == 1 && v.nrot[v.bnode[nbr[s]]] == 1) en += -logf(expf(-val[s]));
    const int nbp = v.stats[4];
    for (int e = c.gtid; e < nbp; e += c.gsize) {
        int f = v.bpf[e], s = v.bps[e];
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                float p = v.marg[(i * 6 + j) * cap + e];
                float q = v.P[e * 36 + i * 6 + j] * v.ncur[f * 6 + i] * v.ncur[s * 6 + j];
                en += p * logf((1e-10f + p) / (1e-10f + q));
            }
    }
    return team_sum(c, en);
}

// Edge sensitivities of bead pairs (src/rotamer.cpp:1057-1066) and the 1-body
// sensitivities (:1074-1085).
__device__ void rot_sens(Ctx& c, const PairGraph& G, const RotamerDev& R) {
    RotView v = rot_view(c, R);
    const int* row = c.I + G.row;
    const int* nbr = c.I + G.nbr;
    const int* act = c.I + G.act;
    float* es = c.A + G.esens;
    const int* loc1 = c.CI + G.loc1;
    const int E = c.I[G.state + 3];
    const int cap = v.cap;
    for (int s = c.gtid; s < E; s += c.gsize) {
        if (!act[s]) { es[s] = 0.f; continue; }
        int i = row[s], j = nbr[s];
        int a = v.bnode[i], b = v.bnode[j];
        int na = v.nrot[a], nb = v.nrot[b];
        int ra = v.brot[i], rb = v.brot[j];
        float sv;
        if (na == 1 && nb == 1) sv = 1.f;
        else if (na == 1) sv = v.ncur[b * 6 + rb];
        else if (nb == 1) sv = v.ncur[a * 6 + ra];
        else {
            int e = v.eidx[v.scand[s]];
            int q = na > nb ? rb * 6 + ra : ra * 6 + rb;
            sv = v.marg[q * cap + e];
        }
        es[s] = sv;
    }
    for (int b = c.gtid; b < G.n1; b += c.gsize) {
        float m = v.ncur[v.bnode[b] * 6 + v.brot[b]];
        for (int q = 0; q < R.n_prob; ++q) c.A[R.prob_sens[q] + loc1[b] * R.prob_w[q]] += m;
    }
    team_sync(c);
}

}  // namespace upside
