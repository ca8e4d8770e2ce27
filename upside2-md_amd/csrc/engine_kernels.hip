// Persistent force-and-integration kernel of the MI355X Upside engine.
//
// One workgroup evaluates one system: it interprets the lowered node program
// (value pass in germ order, sensitivity pass in deriv order, as
// DerivEngine::compute, src/deriv_engine.cpp:239-260), then applies the
// symplectic-Euler stages of integration_cycle (src/deriv_engine.cpp:317-331)
// and the Ornstein-Uhlenbeck thermostat (src/thermostat.cpp:9-18), for many
// rounds without returning to the host.  Each op cites the reference node it
// reproduces.
#include "dev_rotamer.h"

namespace upside {

// operand pointers, typed as global memory (see gmem)
#define OUT(op) gmem(c.A + (op).out)
#define SENS(op) gmem(c.A + (op).sens)
#define IN(op, k) gmem(c.A + (op).in_out[k])
#define INS(op, k) gmem(c.A + (op).in_sens[k])
#define CIA(op, k) gmem(c.CI + (op).ia[k])
#define CFA(op, k) gmem(c.CF + (op).ia[k])

// Profile accumulation (op slots by index, fixed slots through pacc).
__device__ __forceinline__ void prof_op(const Ctx& c, int slot, unsigned long long t0) {
    if (c.prof && c.gtid == 0)
        __hip_atomic_fetch_add(c.prof + slot, wall_clock64() - t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void put_pot(Ctx& c, const SysDesc& S, const OpRec& op, float e) {
    if (c.gtid == 0) c.A[S.pot + op.pot_slot] = e;
}

// ---- Distance3D (src/coord_basic.cpp:13-66) --------------------------------
__device__ void dist3d_fwd(Ctx& c, const OpRec& op) {
    auto id = CIA(op, 0);
    auto der = gmem(c.A + op.scratch);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto a = IN(op, 0) + id[2 * e] * op.in_w[0];
        auto b = IN(op, 1) + id[2 * e + 1] * op.in_w[1];
        float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        float m2 = dx * dx + dy * dy + dz * dz;
        float d = sqrtf(m2);
        OUT(op)[e] = d;
        float inv = d != 0.f ? rsqrtf(m2) : 0.f;
        der[e * 3 + 0] = dx * inv; der[e * 3 + 1] = dy * inv; der[e * 3 + 2] = dz * inv;
    }
}
__device__ void dist3d_bwd(Ctx& c, const OpRec& op) {
    auto der = gmem(c.A + op.scratch);
    auto con = gmem(c.A + op.ia[1]);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        float s = SENS(op)[e];
        for (int k = 0; k < 3; ++k) {
            con[(e * 2 + 0) * 3 + k] = s * der[e * 3 + k];
            con[(e * 2 + 1) * 3 + k] = -s * der[e * 3 + k];
        }
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
    plan_gather(c, op.plans[1]);
}

// ---- Angle: cosine at the third atom (src/coord_basic.cpp:183-243) ---------
__device__ void angle_fwd(Ctx& c, const OpRec& op) {
    auto id = CIA(op, 0);
    auto der = gmem(c.A + op.scratch);
    auto P = IN(op, 0);
    int w = op.in_w[0];
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto a1 = P + id[3 * e] * w;
        auto a2 = P + id[3 * e + 1] * w;
        auto a3 = P + id[3 * e + 2] * w;
        float x1[3], x2[3];
        for (int k = 0; k < 3; ++k) { x1[k] = a1[k] - a3[k]; x2[k] = a2[k] - a3[k]; }
        float i1 = rsqrtf(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
        float i2 = rsqrtf(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
        float h1[3], h2[3];
        for (int k = 0; k < 3; ++k) { h1[k] = x1[k] * i1; h2[k] = x2[k] * i2; }
        float dp = h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2];
        OUT(op)[e] = dp;
        for (int k = 0; k < 3; ++k) {
            der[e * 6 + k] = (h2[k] - h1[k] * dp) * i1;
            der[e * 6 + 3 + k] = (h1[k] - h2[k] * dp) * i2;
        }
    }
}
__device__ void angle_bwd(Ctx& c, const OpRec& op) {
    auto der = gmem(c.A + op.scratch);
    auto con = gmem(c.A + op.ia[1]);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        float s = SENS(op)[e];
        for (int k = 0; k < 3; ++k) {
            float d1 = s * der[e * 6 + k], d2 = s * der[e * 6 + 3 + k];
            con[(e * 3 + 0) * 3 + k] = d1;
            con[(e * 3 + 1) * 3 + k] = d2;
            con[(e * 3 + 2) * 3 + k] = -d1 - d2;
        }
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
}

// ---- Dihedral (src/coord_basic.cpp:311-374, src/vector_math.h:764-797) -----
template <class P, class Q, class O>
__device__ __forceinline__ void cross3(P a, Q b, O o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
template <class P, class Q>
__device__ __forceinline__ float dot3(P a, Q b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ void dihedral_fwd(Ctx& c, const OpRec& op) {
    auto id = CIA(op, 0);
    auto jac = gmem(c.A + op.scratch);
    auto P = IN(op, 0);
    int w = op.in_w[0];
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto q = id + 4 * e;
        if (q[0] < 0 || q[3] < 0) {
            OUT(op)[e] = -1.3963f;
            for (int k = 0; k < 12; ++k) jac[e * 12 + k] = 0.f;
            continue;
        }
        auto r1 = P + q[0] * w; auto r2 = P + q[1] * w; auto r3 = P + q[2] * w; auto r4 = P + q[3] * w;
        float F[3], G[3], H[3], A[3], B[3], C[3];
        for (int k = 0; k < 3; ++k) { F[k] = r1[k] - r2[k]; G[k] = r2[k] - r3[k]; H[k] = r4[k] - r3[k]; }
        cross3(F, G, A);
        cross3(H, G, B);
        cross3(B, A, C);
        float iA2 = 1.f / dot3(A, A), iB2 = 1.f / dot3(B, B);
        float G2 = dot3(G, G);
        float iG = rsqrtf(G2);
        float Gm = G2 * iG;
        float fg = dot3(F, G) * iA2 * iG, hg = dot3(H, G) * iB2 * iG;
        for (int k = 0; k < 3; ++k) {
            float d1 = -Gm * iA2 * A[k];
            float d4 = Gm * iB2 * B[k];
            float fm = fg * A[k] - hg * B[k];
            jac[e * 12 + k] = d1;
            jac[e * 12 + 3 + k] = -d1 + fm;
            jac[e * 12 + 6 + k] = -d4 - fm;
            jac[e * 12 + 9 + k] = d4;
        }
        OUT(op)[e] = atan2f(dot3(C, G), dot3(A, B) * Gm);
    }
}
__device__ void dihedral_bwd(Ctx& c, const OpRec& op) {
    auto jac = gmem(c.A + op.scratch);
    auto con = gmem(c.A + op.ia[1]);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        float s = SENS(op)[e];
        for (int k = 0; k < 12; ++k) con[e * 12 + k] = s * jac[e * 12 + k];
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
}

// ---- Spring (src/spring.cpp:9-81) -------------------------------------------
__device__ void spring_fwd(Ctx& c, const SysDesc& S, const OpRec& op) {
    auto id = CIA(op, 0);
    auto eq = CFA(op, 1);
    auto kk = CFA(op, 2);
    int dim1 = op.ia[3], pbc = op.ia[4], n = op.ia[5];
    float box = op.fa[0];
    auto con = gmem(c.A + op.scratch);
    float pot = 0.f;
    for (int e = c.gtid; e < n; e += c.gsize) {
        float x = IN(op, 0)[id[e] * op.in_w[0] + dim1];
        float ex = x - eq[e];
        float s2 = ex * ex;
        if (pbc && ex != 0.f) {
            float e1 = x - eq[e] - box, e2 = x - eq[e] + box;
            if (e1 * e1 < s2) { s2 = e1 * e1; ex = e1; }
            if (e2 * e2 < s2) { s2 = e2 * e2; ex = e2; }
        }
        pot += 0.5f * kk[e] * s2;
        con[e] = kk[e] * ex;
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
    plan_gather(c, op.plans[0]);
}

// ---- rama_coord / cat_pos (src/coord_map.cpp:69-161) -----------------------
__device__ void rama_coord_fwd(Ctx& c, const OpRec& op) {
    auto i1 = CIA(op, 0);
    auto i2 = CIA(op, 1);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        OUT(op)[2 * e] = IN(op, 0)[i1[e] * op.in_w[0]];
        OUT(op)[2 * e + 1] = IN(op, 1)[i2[e] * op.in_w[1]];
    }
}
__device__ void cat_pos_fwd(Ctx& c, const OpRec& op) {
    auto idx = CIA(op, 0);
    int n1 = op.ia[1], w = op.width;
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        int k = e < n1 ? 0 : 1;
        auto src = IN(op, k) + idx[e] * op.in_w[k];
        for (int d = 0; d < w; ++d) OUT(op)[e * w + d] = src[d];
    }
}
__device__ void passthrough_bwd(Ctx& c, const OpRec& op) {
    plan_gather(c, op.plans[0]);
    plan_gather(c, op.plans[1]);
}

// ---- rama_map_pot (src/rama_map_pot.cpp:15-94) -------------------------------
__device__ void rama_map_fwd(Ctx& c, const SysDesc& S, const OpRec& op) {
    auto res = CIA(op, 0);
    auto mid = CIA(op, 1);
    auto coef = CFA(op, 2);
    int nx = op.ia[3], n = op.ia[4];
    float scale = op.fa[0], shift = op.fa[1];
    auto con = gmem(c.A + op.scratch);
    float pot = 0.f;
    for (int e = c.gtid; e < n; e += c.gsize) {
        auto r = IN(op, 0) + res[e] * op.in_w[0];
        float x = (r[0] + shift) * scale, y = (r[1] + shift) * scale;
        int xb = (int)x, yb = (int)y;
        xb = min(max(xb, 0), nx - 1); yb = min(max(yb, 0), nx - 1);
        float v, dx, dy;
        bicubic(coef + ((size_t)(mid[e] * nx + xb) * nx + yb) * 16, x - xb, y - yb, v, dx, dy);
        pot += v;
        c.A[op.scratch + 2 * n + e] = v;  // per-residue potential (logging)
        con[2 * e] = dx * scale;
        con[2 * e + 1] = dy * scale;
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
    plan_gather(c, op.plans[0]);
}

// ---- affine_alignment (src/eig.cpp:277-471) ----------------------------------
// 4x4 symmetric eigenproblem per residue: Householder tridiagonalisation then
// implicit QR with Wilkinson shift (Golub & Van Loan 8.3), in groups of four
// residues that iterate together as the reference's SIMD lanes do.
__device__ __forceinline__ int pk(int i, int j) {
    int a = i < j ? i : j, b = i < j ? j : i;
    return a == 0 ? b : a == 1 ? 3 + b : a == 2 ? 5 + b : 9;
}
__device__ __forceinline__ bool quad_any(bool p) {
    uint64_t m = __ballot(p);
    int lane = threadIdx.x & 63;
    return ((m >> (lane & ~3)) & 0xFull) != 0;
}

__device__ void eig4_grouped(float* A, float* d, float* rot) {
    float beta[2];
    for (int k = 0; k < 2; ++k) {
        int m = 3 - k;
        float x0 = A[pk(k, k + 1)];
        float sigma2 = 1e-20f;
        for (int i = 1; i < m; ++i) { float xi = A[pk(k, k + 1 + i)]; sigma2 += xi * xi; }
        float mu = sqrtf(x0 * x0 + sigma2);
        float s = 0.f < x0 ? -sigma2 / (x0 + mu) : x0 - mu;
        float b = 2.f * s * s / (sigma2 + s * s);
        A[pk(k, k + 1)] = mu;
        float is = 1.f / s;
        for (int i = 1; i < m; ++i) A[pk(k, k + 1 + i)] *= is;
        beta[k] = b;
        float v[3], p[3], w[3];
        for (int j = 0; j < m; ++j) v[j] = j == 0 ? 1.f : A[pk(k, k + 1 + j)];
        for (int i = 0; i < m; ++i) {
            float acc = 0.f;
            for (int j = 0; j < m; ++j) acc += A[pk(i + k + 1, j + k + 1)] * v[j];
            p[i] = acc * b;
        }
        float pdv = 0.f;
        for (int i = 0; i < m; ++i) pdv += p[i] * v[i];
        for (int i = 0; i < m; ++i) w[i] = p[i] - (0.5f * b * pdv) * v[i];
        for (int i = 0; i < m; ++i)
            for (int j = i; j < m; ++j) A[pk(i + k + 1, j + k + 1)] -= v[i] * w[j] + v[j] * w[i];
    }
    float u[3];
    for (int i = 0; i < 4; ++i) d[i] = A[pk(i, i)];
    for (int i = 0; i < 3; ++i) u[i] = A[pk(i, i + 1)];
    for (int i = 0; i < 16; ++i) rot[i] = (i % 5 == 0) ? 1.f : 0.f;
    float v0[3] = {1.f, A[pk(0, 2)], A[pk(0, 3)]};
    float v1[2] = {1.f, A[pk(1, 3)]};
    for (int i = 1; i < 4; ++i)
        for (int j = i; j < 4; ++j) rot[i * 4 + j] -= beta[0] * (v0[i - 1] * v0[j - 1]);
    for (int i = 2; i < 4; ++i)
        for (int j = i; j < 4; ++j) rot[i * 4 + j] -= beta[1] * (v1[i - 2] * v1[j - 2]);
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) rot[j * 4 + i] = rot[i * 4 + j];
    float coeff = beta[0] * beta[1] * (v0[1] * v1[0] + v0[2] * v1[1]);
    for (int i = 2; i < 4; ++i)
        for (int j = 1; j < 4; ++j) rot[i * 4 + j] += coeff * v1[i - 2] * v0[j - 1];
    const float tol = 1e-5f;
    for (int it = 0; it < 100; ++it) {
        for (int i = 0; i < 3; ++i)
            if (fabsf(u[i]) <= tol * (fabsf(d[i]) + fabsf(d[i + 1]))) u[i] = 0.f;
        int q;
        if (quad_any(u[2] != 0.f)) q = 0;
        else if (quad_any(u[1] != 0.f)) q = 1;
        else if (quad_any(u[0] != 0.f)) q = 2;
        else break;
        int p = 4 - q - 1;
        for (; p > 0; --p)
            if (!quad_any(u[p - 1] != 0.f)) break;
        int n = 4 - q - p;
        auto dd = d + p;
        auto uu = u + p;
        auto rr = rot + 4 * p;
        float dval = 0.5f * (dd[n - 2] - dd[n - 1]) + 1e-20f;
        float un = uu[n - 2];
        float mu = dd[n - 1] - un * un / (dval + copysignf(sqrtf(dval * dval + un * un), dval));
        float x = dd[0] - mu, z = uu[0];
        for (int k = 0; k < n - 1; ++k) {
            float cc, ss;
            if (z == 0.f) { cc = 1.f; ss = 0.f; }
            else { float ir = rsqrtf(x * x + z * z); cc = x * ir; ss = -z * ir; }
            for (int j = 0; j < 4; ++j) {
                float t1 = rr[k * 4 + j], t2 = rr[(k + 1) * 4 + j];
                rr[k * 4 + j] = cc * t1 - ss * t2;
                rr[(k + 1) * 4 + j] = ss * t1 + cc * t2;
            }
            if (k > 0) uu[k - 1] = cc * x - ss * z;
            float T00 = dd[k], T11 = dd[k + 1], T01 = uu[k];
            dd[k] = T00 * cc * cc - T01 * 2.f * cc * ss + T11 * ss * ss;
            dd[k + 1] = T00 * ss * ss + T01 * 2.f * cc * ss + T11 * cc * cc;
            uu[k] = (T00 - T11) * cc * ss + T01 * (cc * cc - ss * ss);
            x = uu[k];
            if (k < n - 2) { z = -uu[k + 1] * ss; uu[k + 1] *= cc; }
        }
    }
    for (int i = 1; i < 4; ++i) {
        if (d[0] < d[i]) {
            float t = d[0]; d[0] = d[i]; d[i] = t;
            for (int j = 0; j < 4; ++j) { float r = rot[j]; rot[j] = rot[i * 4 + j]; rot[i * 4 + j] = r; }
        }
    }
}

__device__ void affine_fwd(Ctx& c, const OpRec& op) {
    auto atoms = CIA(op, 0);
    auto ref = CFA(op, 1);
    auto ev = gmem(c.A + op.scratch);            // n*4 eigenvalues
    auto ec = gmem(c.A + op.scratch + 4 * op.n_elem);  // n*16 eigenvector rows
    auto P = IN(op, 0);
    int w = op.in_w[0];
    int n = op.n_elem;
    for (int base = 0; base < n; base += c.gsize) {
        int r = base + c.gtid;
        int rr = r < n ? r : (r & ~3);       // padding lanes repeat lane 0 of their quad
        if (rr >= n) continue;              // whole quad past the end
        float x[3][3], center[3];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) x[a][k] = P[atoms[rr * 3 + a] * w + k];
        for (int k = 0; k < 3; ++k) center[k] = (1.f / 3.f) * (x[0][k] + x[1][k] + x[2][k]);
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) x[a][k] -= center[k];
        auto g = ref + rr * 9;
        float R[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[i][j] = x[0][j] * g[i] + x[1][j] * g[3 + i] + x[2][j] * g[6 + i];
        float F[10] = {R[0][0] + R[1][1] + R[2][2], R[1][2] - R[2][1], R[2][0] - R[0][2], R[0][1] - R[1][0],
                       R[0][0] - R[1][1] - R[2][2], R[0][1] + R[1][0], R[0][2] + R[2][0],
                       -R[0][0] + R[1][1] - R[2][2], R[1][2] + R[2][1], -R[0][0] - R[1][1] + R[2][2]};
        float d[4], rot[16];
        eig4_grouped(F, d, rot);
        if (r >= n) continue;
        for (int k = 0; k < 4; ++k) ev[r * 4 + k] = d[k];
        for (int k = 0; k < 16; ++k) ec[r * 16 + k] = rot[k];
        auto o = OUT(op) + r * 7;
        o[0] = center[0]; o[1] = center[1]; o[2] = center[2];
        o[3] = rot[0]; o[4] = rot[1]; o[5] = rot[2]; o[6] = rot[3];
    }
}
__device__ void affine_bwd(Ctx& c, const OpRec& op) {
    auto ref = CFA(op, 1);
    auto ev = gmem(c.A + op.scratch);
    auto ec = gmem(c.A + op.scratch + 4 * op.n_elem);
    auto con = gmem(c.A + op.ia[2]);
    for (int r = c.gtid; r < op.n_elem; r += c.gsize) {
        auto s = SENS(op) + r * 7;
        auto E = ec + r * 16;
        auto q = E;
        float t0 = s[3], t1 = s[4], t2 = s[5];
        float qs[4] = {2.f * (-t0 * q[1] - t1 * q[2] - t2 * q[3]), 2.f * (t0 * q[0] + t1 * q[3] - t2 * q[2]),
                       2.f * (t1 * q[0] + t2 * q[1] - t0 * q[3]), 2.f * (t2 * q[0] + t0 * q[2] - t1 * q[1])};
        float qsd[4], inv[4];
        for (int dd = 0; dd < 4; ++dd) {
            float acc = 0.f;
            for (int i = 0; i < 4; ++i) acc += qs[i] * E[dd * 4 + i];
            qsd[dd] = acc;
        }
        for (int k = 1; k < 4; ++k) inv[k] = 1.f / (ev[r * 4] - ev[r * 4 + k]);
        for (int na = 0; na < 3; ++na) {
            auto g = ref + r * 9 + na * 3;
            float gx = g[0], gy = g[1], gz = g[2];
            // upper triangles (f00 f01 f02 f03 f11 f12 f13 f22 f23 f33), src/eig.cpp:452-467
            float M[3][10] = {{gx, 0.f, gz, -gy, gx, gy, gz, -gx, 0.f, -gx},
                              {gy, -gz, 0.f, gx, -gy, gx, 0.f, gy, gz, -gy},
                              {gz, gy, -gx, 0.f, -gz, 0.f, gx, -gz, gy, gz}};
            for (int j = 0; j < 3; ++j) {
                float dv = (1.f / 3.f) * s[j];
                for (int k = 1; k < 4; ++k) {
                    float t = 0.f;
                    for (int a = 0; a < 4; ++a)
                        for (int b = a; b < 4; ++b) {
                            float fab = M[j][pk(a, b)];
                            float term = a == b ? E[k * 4 + a] * E[b] : E[k * 4 + a] * E[b] + E[k * 4 + b] * E[a];
                            t += fab * term;
                        }
                    dv += inv[k] * t * qsd[k];
                }
                con[(r * 3 + na) * 3 + j] = dv;
            }
        }
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
}

// ---- placement (src/placement.cpp:20-314) -------------------------------------
__device__ void placement_fwd(Ctx& c, const OpRec& op) {
    auto ares = CIA(op, 0);
    auto layer = CIA(op, 1);
    auto data = CFA(op, 2);
    int is_rama = op.ia[3], dim = op.ia[7];
    auto aff = IN(op, 0);
    int wa = op.in_w[0];
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto o = OUT(op) + e * op.width;
        if (is_rama) {
            auto rres = CIA(op, 4);
            int nx = op.ia[5], ny = op.ia[6];
            auto r = IN(op, 1) + rres[e] * op.in_w[1];
            float x = (r[0] + op.fa[2]) * op.fa[0], y = (r[1] + op.fa[2]) * op.fa[1];
            int xb = min(max((int)x, 0), nx - 1), yb = min(max((int)y, 0), ny - 1);
            float v, dx, dy;
            bicubic(data + ((size_t)(layer[e] * nx + xb) * ny + yb) * 16, x - xb, y - yb, v, dx, dy);
            o[0] = v;
            c.A[op.scratch + 2 * e] = dx * op.fa[0];
            c.A[op.scratch + 2 * e + 1] = dy * op.fa[1];
            continue;
        }
        auto a = aff + ares[e] * wa;
        float U[9];
        quat_to_rot(a + 3, U);
        auto v = data + layer[e] * dim;
        for (int k = 0; k < 3; ++k) {
            o[k] = U[3 * k] * v[0] + U[3 * k + 1] * v[1] + U[3 * k + 2] * v[2] + a[k];
            o[3 + k] = U[3 * k] * v[3] + U[3 * k + 1] * v[4] + U[3 * k + 2] * v[5];
        }
        if (dim == 7) o[6] = v[6];
    }
}
__device__ void placement_bwd(Ctx& c, const OpRec& op) {
    int is_rama = op.ia[3];
    auto con = gmem(c.A + op.ia[8]);
    if (is_rama) {
        for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
            float s = SENS(op)[e];
            con[2 * e] = s * c.A[op.scratch + 2 * e];
            con[2 * e + 1] = s * c.A[op.scratch + 2 * e + 1];
        }
        team_sync(c);
        plan_gather(c, op.plans[1]);
        return;
    }
    auto ares = CIA(op, 0);
    auto aff = IN(op, 0);
    int wa = op.in_w[0];
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto s = SENS(op) + e * op.width;
        auto x = OUT(op) + e * op.width;
        auto t = aff + ares[e] * wa;
        float p[3] = {x[0] - t[0], x[1] - t[1], x[2] - t[2]};
        float tq1[3], tq2[3];
        cross3(p, s, tq1);
        cross3(x + 3, s + 3, tq2);
        for (int k = 0; k < 3; ++k) {
            con[e * 6 + k] = s[k];
            con[e * 6 + 3 + k] = tq1[k] + tq2[k];
        }
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
}

// ---- infer_H_O (src/hbond.cpp:20-127) ------------------------------------------
__device__ void infer_fwd(Ctx& c, const OpRec& op) {
    auto id = CIA(op, 0);
    auto bl = CFA(op, 1);
    auto cache = gmem(c.A + op.scratch);
    auto P = IN(op, 0);
    int w = op.in_w[0];
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto pc = P + id[3 * e] * w; auto cc = P + id[3 * e + 1] * w; auto nc = P + id[3 * e + 2] * w;
        float pr[3], nx[3], ds[3];
        for (int k = 0; k < 3; ++k) { pr[k] = pc[k] - cc[k]; nx[k] = nc[k] - cc[k]; }
        float pi = rsqrtf(dot3(pr, pr)), ni = rsqrtf(dot3(nx, nx));
        for (int k = 0; k < 3; ++k) { pr[k] *= pi; nx[k] *= ni; ds[k] = pr[k] + nx[k]; }
        float di = rsqrtf(dot3(ds, ds));
        for (int k = 0; k < 3; ++k) ds[k] *= di;
        auto cs = cache + e * 12;
        for (int k = 0; k < 3; ++k) { cs[k] = pr[k]; cs[4 + k] = nx[k]; cs[8 + k] = ds[k]; }
        cs[3] = pi; cs[7] = ni; cs[11] = di;
        auto o = OUT(op) + e * 6;
        for (int k = 0; k < 3; ++k) { o[k] = cc[k] - bl[e] * ds[k]; o[3 + k] = -ds[k]; }
    }
}
__device__ void infer_bwd(Ctx& c, const OpRec& op) {
    auto bl = CFA(op, 1);
    auto cache = gmem(c.A + op.scratch);
    auto con = gmem(c.A + op.ia[2]);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto s = SENS(op) + e * 6;
        auto cs = cache + e * 12;
        float sn[3], sd[3], sp[3], sx[3];
        for (int k = 0; k < 3; ++k) sn[k] = s[3 + k] + bl[e] * s[k];
        float t = dot3(cs + 8, sn);
        for (int k = 0; k < 3; ++k) sd[k] = cs[11] * (t * cs[8 + k] - sn[k]);
        float tp = dot3(cs, sd), tn = dot3(cs + 4, sd);
        for (int k = 0; k < 3; ++k) {
            sp[k] = -cs[3] * (tp * cs[k] - sd[k]);
            sx[k] = -cs[7] * (tn * cs[4 + k] - sd[k]);
        }
        for (int k = 0; k < 3; ++k) {
            con[(e * 3 + 0) * 3 + k] = sp[k];
            con[(e * 3 + 1) * 3 + k] = s[k] - sp[k] - sx[k];
            con[(e * 3 + 2) * 3 + k] = sx[k];
        }
    }
    team_sync(c);
    plan_gather(c, op.plans[0]);
}

// A pair-graph capacity overflow voids the evaluation: the system's error
// word is set, the remaining ops are skipped and the host grows the
// capacities and repeats the run.
__device__ void graph_overflow(Ctx& c, const SysDesc& S) {
    if (c.gtid == 0) c.I[c.err] = 1;
    team_sync(c);
}

// ---- protein_hbond (src/hbond.cpp:357-435) ---------------------------------------
__device__ void protein_hbond_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G) {
    if (!pg_refresh(c, G, op)) { graph_overflow(c, S); return; }
    int nd = G.n1;
    auto loc1 = gmem(c.CI + G.loc1);
    auto loc2 = gmem(c.CI + G.loc2);
    auto ss = gmem(c.A + op.scratch);  // per-site sum of hb_log
    pg_row_sums(c, G, false, ss, 1);
    pg_row_sums(c, G, true, ss + nd, 1);
    team_sync(c);
    for (int v = c.gtid; v < op.n_elem; v += c.gsize) {
        bool don = v < nd;
        int i = don ? v : v - nd;
        auto src = IN(op, 0) + (don ? loc1[i] : loc2[i]) * op.in_w[0];
        auto o = OUT(op) + v * 7;
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = src[k];
        o[6] = 1.f - expf(-ss[v]);
    }
}
__device__ void protein_hbond_bwd(Ctx& c, const OpRec& op, const PairGraph& G) {
    int nd = G.n1;
    auto ss = gmem(c.A + op.scratch);
    for (int v = c.gtid; v < op.n_elem; v += c.gsize) ss[v] = SENS(op)[v * 7 + 6] * (1.f - OUT(op)[v * 7 + 6]);
    team_sync(c);
    auto row = gmem(c.I + G.row);
    auto nbr = gmem(c.I + G.nbr);
    auto es = gmem(c.A + G.esens);
    const int E = c.I[G.state + 3];
    for (int s = c.gtid; s < E; s += c.gsize) es[s] = ss[row[s]] + ss[nd + nbr[s]];
    team_sync(c);
    pg_gather(c, G);
    // pass the coordinate sensitivities through to infer_H_O (:419-433)
    auto loc1 = gmem(c.CI + G.loc1);
    auto loc2 = gmem(c.CI + G.loc2);
    for (int i = c.gtid; i < G.n1; i += c.gsize)
#pragma unroll
        for (int k = 0; k < 6; ++k) INS(op, 0)[loc1[i] * op.in_w[0] + k] += SENS(op)[i * 7 + k];
    team_sync(c);
    for (int j = c.gtid; j < G.n2; j += c.gsize)
#pragma unroll
        for (int k = 0; k < 6; ++k) INS(op, 0)[loc2[j] * op.in_w[0] + k] += SENS(op)[(nd + j) * 7 + k];
    team_sync(c);
}

// ---- hbond_coverage (src/hbond.cpp:438-484) ----------------------------------------
__device__ void hbond_coverage_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G) {
    if (!pg_refresh(c, G, op)) { graph_overflow(c, S); return; }
    pg_row_sums(c, G, true, OUT(op), 1);
}
__device__ void hbond_coverage_bwd(Ctx& c, const OpRec& op, const PairGraph& G) {
    auto nbr = gmem(c.I + G.nbr);
    auto es = gmem(c.A + G.esens);
    const int E = c.I[G.state + 3];
    for (int s = c.gtid; s < E; s += c.gsize) es[s] = SENS(op)[nbr[s]];
    team_sync(c);
    pg_gather(c, G);
}

// ---- hbond_energy (src/hbond.cpp:487-739) -------------------------------------------
// ia: 0 n_res, 1 hb1 off (n_res+1), 2 hb1 idx, 3 hb2 off (n_res), 4 hb2 idx,
//     5 first donor of residue, 6 first acceptor of residue (virtual index or -1)
__device__ void hbond_energy_fwd(Ctx& c, const SysDesc& S, const OpRec& op) {
    int n_res = op.ia[0];
    auto h1o = CIA(op, 1);
    auto h1i = CIA(op, 2);
    auto h2o = CIA(op, 3);
    auto h2i = CIA(op, 4);
    auto fd = CIA(op, 5);
    auto fa = CIA(op, 6);
    auto pp = IN(op, 0);
    auto pps = INS(op, 0);
    auto rama = IN(op, 1);
    auto ramas = INS(op, 1);
    int wp = op.in_w[0], wr = op.in_w[1];
    auto P = op.fa;
    float pot = 0.f;
    for (int i = c.gtid; i < n_res; i += c.gsize) {
        float phi = rama[i * wr], psi = rama[i * wr + 1];
        float p1, d1, p2, d2, s1, e1, s2, e2;
        compact_sigmoid(phi - P[4], P[5], p1, d1);
        compact_sigmoid(phi - P[6], P[7], p2, d2);
        compact_sigmoid(psi - P[8], P[9], s1, e1);
        compact_sigmoid(psi - P[10], P[11], s2, e2);
        float turn = (float)((1. - (double)p1) * p2);
        float psis = (float)((1. - (double)s1) * s2);
        float nturn = (float)(1. - (double)turn), npsi = (float)(1. - (double)psis);
        float helix = nturn * psis, sheet = nturn * npsi;
        float dturn = d2 - (d1 * p2 + p1 * d2);
        float dpsis = e2 - (e1 * s2 + s1 * e2);
        float Eh = P[0] * helix + P[1] * sheet + P[2] * turn;
        float dphi = P[0] * (-dturn * psis) + P[1] * (-dturn * npsi) + P[2] * dturn;
        float dpsi = P[0] * (nturn * dpsis) + P[1] * (nturn * -dpsis);
        float hb = 0.f;
        for (int k = h1o[i]; k < h1o[i + 1]; ++k) hb += pp[h1i[k] * wp + 6];
        if (hb == 0.f) continue;
        pot += hb * Eh;
        if (fd[i] >= 0) pps[fd[i] * wp + 6] += Eh;
        if (fa[i] >= 0) pps[fa[i] * wp + 6] += Eh;
        ramas[i * wr] += hb * dphi;
        ramas[i * wr + 1] += hb * dpsi;
    }
    team_sync(c);
    for (int i = c.gtid; i < n_res - 1; i += c.gsize) {
        float hb2 = 0.f;
        for (int k = h2o[i]; k < h2o[i + 1]; ++k) hb2 += pp[h2i[k] * wp + 6];
        if (hb2 <= 1.f) continue;
        double v, dv;
        compact_sigmoid(1.5 - (double)hb2, 2.0, v, dv);
        pot += (float)(v * P[3]);
        float dE = (float)(-dv * P[3]);
        if (fd[i + 1] >= 0) pps[fd[i + 1] * wp + 6] += dE;
        if (fa[i] >= 0) pps[fa[i] * wp + 6] += dE;
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
}

// ---- weighted_pos (src/environment.cpp:489-533) -------------------------------------
__device__ void weighted_pos_fwd(Ctx& c, const OpRec& op) {
    auto ip = CIA(op, 0);
    auto iw = CIA(op, 1);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) {
        auto p = IN(op, 0) + ip[e] * op.in_w[0];
        auto o = OUT(op) + e * 4;
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        o[3] = expf(-IN(op, 1)[iw[e] * op.in_w[1]]);
    }
}
__device__ void weighted_pos_bwd(Ctx& c, const OpRec& op) {
    auto con = gmem(c.A + op.scratch);
    for (int e = c.gtid; e < op.n_elem; e += c.gsize) con[e] = -OUT(op)[e * 4 + 3] * SENS(op)[e * 4 + 3];
    team_sync(c);
    plan_gather(c, op.plans[0]);
    plan_gather(c, op.plans[1]);
}

// ---- environment_coverage (src/environment.cpp:288-348) -----------------------------
// Output row i holds the coverage of element i by each residue type: the
// active edge values of its slots binned by the partner's amino-acid type.
__device__ void env_coverage_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G) {
    if (!pg_refresh(c, G, op)) { graph_overflow(c, S); return; }
    int naa = op.ia[0];
    auto aa = CIA(op, 1);
    auto id2 = gmem(c.CI + G.id2);
    auto off = gmem(c.I + G.off);
    auto nbr = gmem(c.I + G.nbr);
    auto val = gmem(c.A + G.val);
    auto act = gmem(c.I + G.act);
    auto out = OUT(op);
    // one 16-lane group per element: each lane strides the element's slots
    // into its own type bins (LDS), then lane t of the group sums bin t over
    // the group's lanes in lane order
    lp<float> mine = lmem(c.lds) + c.tid * naa;
    lp<float> grp_bins = lmem(c.lds) + (c.tid & ~(GW - 1)) * naa;
    const int grp = c.gtid / GW, gl = c.gtid % GW, NG = c.gsize / GW;
    for (int base = 0; base < G.n1; base += NG) {
        const int i = base + grp;
        const bool ok = i < G.n1;
        for (int t = 0; t < naa; ++t) mine[t] = 0.f;
        const int e = ok ? off[i + 1] : 0;
        for (int s = (ok ? off[i] : 0) + gl; s < e; s += GW)
            if (act[s]) mine[aa[id2[nbr[s]]]] += val[s];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int t = gl; t < naa; t += GW) {
            float acc = 0.f;
            for (int l = 0; l < GW; ++l) acc += grp_bins[l * naa + t];
            if (ok) out[i * naa + t] = acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // bins are cleared for the next element only after every lane read them
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}
__device__ void env_coverage_bwd(Ctx& c, const OpRec& op, const PairGraph& G) {
    int naa = op.ia[0];
    auto aa = CIA(op, 1);
    auto id2 = gmem(c.CI + G.id2);
    auto row = gmem(c.I + G.row);
    auto nbr = gmem(c.I + G.nbr);
    auto es = gmem(c.A + G.esens);
    const int E = c.I[G.state + 3];
    for (int s = c.gtid; s < E; s += c.gsize) es[s] = SENS(op)[row[s] * naa + aa[id2[nbr[s]]]];
    team_sync(c);
    pg_gather(c, G);
}

// ---- hb_environment_coverage (src/environment.cpp:350-385) ------------------------
__device__ void hb_env_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G) {
    if (!pg_refresh(c, G, op)) { graph_overflow(c, S); return; }
    pg_row_sums(c, G, false, OUT(op), 1);
}
__device__ void hb_env_bwd(Ctx& c, const OpRec& op, const PairGraph& G) {
    auto row = gmem(c.I + G.row);
    auto es = gmem(c.A + G.esens);
    const int E = c.I[G.state + 3];
    for (int s = c.gtid; s < E; s += c.gsize) es[s] = SENS(op)[row[s]];
    team_sync(c);
    pg_gather(c, G);
}

// ---- sigmoid_coupling (src/environment.cpp:851-1003) -------------------------------
// ia: 0 n_res, 1 coupling types, 2 weights (nt*nt), 3 scale, 4 center, 5 sharpness, 6 nt
__device__ void sigmoid_coupling_fwd(Ctx& c, const SysDesc& S, const OpRec& op) {
    int n = op.ia[0], nt = op.ia[6];
    auto ct = CIA(op, 1);
    auto W = CFA(op, 2);
    auto sc = CFA(op, 3);
    auto ce = CFA(op, 4);
    auto sh = CFA(op, 5);
    auto x = IN(op, 0);
    auto xs = INS(op, 0);
    float pot = 0.f;
    for (int r = c.gtid; r < n; r += c.gsize) {
        int t = ct[r];
        float wn = 0.f;
        for (int a = 0; a < nt; ++a) wn += W[t * nt + a] * x[r * nt + a];
        float v, dv;
        compact_sigmoid(wn - ce[t], sh[t], v, dv);
        pot += sc[t] * v;
        for (int a = 0; a < nt; ++a) xs[r * nt + a] += W[t * nt + a] * sc[t] * dv;
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
}

// ---- bb_sigmoid_coupling (src/environment.cpp:1054-1138) ---------------------------
__device__ void bb_sigmoid_fwd(Ctx& c, const SysDesc& S, const OpRec& op) {
    int nt = op.ia[0];
    auto W = CFA(op, 1);
    int n = op.ia[2];
    float scale = op.fa[0], center = op.fa[1], sharp = op.fa[2], hw = op.fa[3];
    auto sc = IN(op, 0);
    auto scs = INS(op, 0);
    auto hb = IN(op, 1);
    auto hbs = INS(op, 1);
    float pot = 0.f;
    for (int r = c.gtid; r < n; r += c.gsize) {
        float bl = 0.f;
        for (int a = 0; a < nt; ++a) bl += W[a] * sc[r * nt + a];
        bl += hw * hb[r * op.in_w[1]];
        float v, dv;
        compact_sigmoid(bl - center, sharp, v, dv);
        pot += scale * v;
        float dp = scale * dv;
        for (int a = 0; a < nt; ++a) scs[r * nt + a] += W[a] * dp;
        hbs[r * op.in_w[1]] += hw * dp;
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
}

// ---- backbone_pairs (src/backbone_steric.cpp:38-146) --------------------------------
__device__ void backbone_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G) {
    if (!pg_refresh(c, G, op)) { graph_overflow(c, S); return; }
    auto val = gmem(c.A + G.val);
    auto act = gmem(c.I + G.act);
    auto es = gmem(c.A + G.esens);
    const int E = c.I[G.state + 3];
    float pot = 0.f;
    for (int s = c.gtid; s < E; s += c.gsize) {
        es[s] = 1.f;
        if (act[s]) pot += val[s];
    }
    pot = team_sum(c, pot);
    put_pot(c, S, op, pot);
    pg_gather(c, G);
}

// ---- rotamer (src/rotamer.cpp:840-850) ----------------------------------------------
// Rotamer node, before the solve: cache check and rebuild, edge energies,
// BP fill.  Returns false on capacity overflow (error word set, team synced).
// The rotamer's pair list (refreshed when the beads moved past half the
// buffer, with the residue-pair candidates) and its edge energies.  Reads only
// the bead placements; false on a capacity overflow.
__device__ bool rotamer_graph(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G, const RotamerDev& R) {
    unsigned long long t0 = ptick(c);
    bool rebuild = pg_pack_and_check(c, G);
    if (rebuild) {
        pg_rebuild(c, G);
        unsigned long long t1 = ptick(c);
        if (c.I[G.state + 0]) rot_rebuild(c, G, R);
        pacc(c, PROF_ROT_REBUILD, t1);
    }
    pacc(c, PROF_REBUILD, t0);
    if (!c.I[G.state + 0]) { graph_overflow(c, S); return false; }
    t0 = ptick(c);
    pg_edges(c, G, op);
    pacc(c, PROF_PG_EDGES, t0);
    return true;
}

__device__ bool rotamer_pre(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G, const RotamerDev& R) {
    if (!c.rot_graph_done && !rotamer_graph(c, S, op, G, R)) return false;
    unsigned long long t0 = ptick(c);
    rot_fill(c, G, R);
    pacc(c, PROF_ROT_FILL, t0);
    return true;
}

// Rotamer node, after the solve (it: iteration count, on gtid 0): statistics,
// edge and 1-body sensitivities, their gather, free energy.  Ends with a team
// barrier.
__device__ void rotamer_post(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G, const RotamerDev& R,
                             int mode, int it) {
    if (c.gtid == 0) {
        auto st = gmem(c.I + R.stats);
        st[0] = it;
        if (it >= R.max_iter - R.chunk - 1) st[1] += 1;
        st[2] += it;
        st[3] += 1;
        st[BP_HIST + min(it, BP_HIST_BINS - 1)] += 1;
    }
    rot_sens(c, G, R);
    pg_gather(c, G);
    float fe = mode ? rot_free_energy(c, G, R) : 0.f;
    put_pot(c, S, op, fe);
}

__device__ void rotamer_fwd(Ctx& c, const SysDesc& S, const OpRec& op, const PairGraph& G,
                            const RotamerDev& R, int mode) {
    if (!rotamer_pre(c, S, op, G, R)) return;
    unsigned long long t0 = ptick(c);
    int it = rot_solve(c, R, c.bp_layout);
    pacc(c, PROF_BP, t0);
    rotamer_post(c, S, op, G, R, mode, it);
}

// ---- program interpreter ----------------------------------------------------------
__device__ void run_value(Ctx& c, const SysDesc& S, const OpRec& op, int mode) {
    const PairGraph* G = op.graph >= 0 ? S.graphs + op.graph : nullptr;
    switch (op.kind) {
    case OP_DIST3D: dist3d_fwd(c, op); break;
    case OP_ANGLE: angle_fwd(c, op); break;
    case OP_DIHEDRAL: dihedral_fwd(c, op); break;
    case OP_SPRING: spring_fwd(c, S, op); break;
    case OP_RAMA_COORD: rama_coord_fwd(c, op); break;
    case OP_CAT_POS: cat_pos_fwd(c, op); break;
    case OP_RAMA_MAP_POT: rama_map_fwd(c, S, op); break;
    case OP_AFFINE: affine_fwd(c, op); break;
    case OP_PLACEMENT: placement_fwd(c, op); break;
    case OP_INFER_HO: infer_fwd(c, op); break;
    case OP_PROTEIN_HBOND: protein_hbond_fwd(c, S, op, *G); break;
    case OP_HBOND_COVERAGE: hbond_coverage_fwd(c, S, op, *G); break;
    case OP_HBOND_ENERGY: hbond_energy_fwd(c, S, op); break;
    case OP_ROTAMER: rotamer_fwd(c, S, op, *G, S.rots[op.rot], mode); return;  // ends with a team barrier
    case OP_WEIGHTED_POS: weighted_pos_fwd(c, op); break;
    case OP_ENV_COVERAGE: env_coverage_fwd(c, S, op, *G); break;
    case OP_HB_ENV_COVERAGE: hb_env_fwd(c, S, op, *G); break;
    case OP_SIGMOID_COUPLING: sigmoid_coupling_fwd(c, S, op); break;
    case OP_BB_SIGMOID_COUPLING: bb_sigmoid_fwd(c, S, op); break;
    case OP_BACKBONE_PAIRS: backbone_fwd(c, S, op, *G); break;
    default: break;
    }
    team_sync(c);
}

__device__ void run_deriv(Ctx& c, const SysDesc& S, const OpRec& op) {
    const PairGraph* G = op.graph >= 0 ? S.graphs + op.graph : nullptr;
    switch (op.kind) {
    case OP_DIST3D: dist3d_bwd(c, op); break;
    case OP_ANGLE: angle_bwd(c, op); break;
    case OP_DIHEDRAL: dihedral_bwd(c, op); break;
    case OP_RAMA_COORD: passthrough_bwd(c, op); break;
    case OP_CAT_POS: passthrough_bwd(c, op); break;
    case OP_AFFINE: affine_bwd(c, op); break;
    case OP_PLACEMENT: placement_bwd(c, op); break;
    case OP_INFER_HO: infer_bwd(c, op); break;
    // these end with their own team barrier (the sensitivity gather)
    case OP_PROTEIN_HBOND: protein_hbond_bwd(c, op, *G); return;
    case OP_HBOND_COVERAGE: hbond_coverage_bwd(c, op, *G); return;
    case OP_WEIGHTED_POS: weighted_pos_bwd(c, op); break;
    case OP_ENV_COVERAGE: env_coverage_bwd(c, op, *G); return;
    case OP_HB_ENV_COVERAGE: hb_env_bwd(c, op, *G); return;
    default: break;  // potential nodes push in their value pass
    }
    team_sync(c);
}

// A team context narrowed to rank 0's workgroup alone (team barriers become
// workgroup barriers, team reductions block reductions).
__device__ __forceinline__ Ctx solo_ctx(const Ctx& c) {
    Ctx l = c;
    l.G = 1;
    l.rank = 0;
    l.gtid = c.tid;
    l.gsize = NT;
    return l;
}

// End of the run of solo ops starting at k (ops [k, end) are all solo).
__device__ __forceinline__ int solo_run_end(const SysDesc& S, int k, int end) {
    while (k < end && S.ops[k].solo) ++k;
    return k;
}

// The deriv ops marked early, in deriv order, on the (sub-)team s that ran the
// rotamer's overlap ops: their sensitivities are complete once those ops are
// done, and nothing the solve or the rotamer's sensitivity push touches is
// read or written by them.  Runs of solo ops go to s's rank 0.
__device__ void run_early_derivs(Ctx& s, const SysDesc& S) {
    const int nd = S.n_germ + S.n_deriv;
    for (int k = S.n_germ; k < nd;) {
        const OpRec& op = S.ops[k];
        if (!op.early) { ++k; continue; }
        s.cur_op = k;
        if (s.G > 1 && op.solo) {
            int k2 = k;
            while (k2 < nd && (!S.ops[k2].early || S.ops[k2].solo)) ++k2;
            if (s.rank == 0) {
                Ctx l = solo_ctx(s);
                for (int q = k; q < k2; ++q) {
                    if (!S.ops[q].early) continue;
                    unsigned long long t0 = ptick(l);
                    run_deriv(l, S, S.ops[q]);
                    prof_op(l, q, t0);
                }
            }
            team_sync(s);
            k = k2;
            continue;
        }
        unsigned long long t0 = ptick(s);
        run_deriv(s, S, op);
        prof_op(s, k, t0);
        ++k;
    }
}

// The ops of one lane of a two-lane section [k0, k1) on half-team s: ops of
// the other lane are skipped (they run on the other half, and the host made
// the two lanes independent: neither reads what the other writes).  A lane's
// solo ops up to its next team op run on the half's rank 0.  Returns false
// when an op set the half's error word.
__device__ bool run_lane(Ctx& s, const SysDesc& S, int k0, int k1, unsigned lanes, int mode) {
    const OpRec& first = S.ops[k0];
    auto mine = [&](const OpRec& o) { return ((lanes >> o.lane) & 1u) != 0u; };
    if (first.rot_graph_op >= 0 && ((lanes >> first.rot_graph_lane) & 1u)) {
        const OpRec& ro = S.ops[first.rot_graph_op];
        s.cur_op = first.rot_graph_op;
        if (!rotamer_graph(s, S, ro, S.graphs[ro.graph], S.rots[ro.rot])) return false;
    }
    for (int k = k0; k < k1;) {
        const OpRec& op = S.ops[k];
        if (!mine(op)) { ++k; continue; }
        s.cur_op = k;
        if (s.G > 1 && op.solo) {
            int k2 = k;
            while (k2 < k1 && (!mine(S.ops[k2]) || S.ops[k2].solo)) ++k2;
            if (s.rank == 0) {
                Ctx l = solo_ctx(s);
                for (int q = k; q < k2; ++q) {
                    if (!mine(S.ops[q])) continue;
                    unsigned long long t0 = ptick(l);
                    run_value(l, S, S.ops[q], mode);
                    prof_op(l, q, t0);
                    if (l.I[l.err]) break;
                }
            }
            team_sync(s);
            if (s.I[s.err]) return false;
            k = k2;
            continue;
        }
        unsigned long long t0 = ptick(s);
        run_value(s, S, op, mode);
        prof_op(s, k, t0);
        if (s.I[s.err]) return false;
        ++k;
    }
    return true;
}

// Value ops [k0, k1) in germ order on the team c.  Ops marked solo
// (elementwise work over at most a few thousand elements) run on rank 0 of
// the team alone: a run of them goes back to back with workgroup barriers and
// ends with one team barrier, since a team barrier costs more than such an
// op.  With SPLIT, a rotamer op with an overlap range runs its solve on rank 0
// while the other ranks evaluate the following independent ops.  Returns
// false when an op set the error word.
template <bool SPLIT>
__device__ bool run_germ_range(Ctx& c, const SysDesc& S, int k0, int k1, int mode) {
    for (int k = k0; k < k1;) {
        const OpRec& op = S.ops[k];
        c.cur_op = k;
        if (SPLIT && op.lane_end > k && op.lane_end <= k1 && c.G >= 4) {
            // section: the team splits into nq sub-teams (four quarters when
            // the program has four lanes, the team eight or more workgroups
            // and the launch asks for them, else two halves), each with its own barrier
            // counter, reduction scratch and error word; sub-team j runs
            // lane j, or lanes j and j+2 (four lanes on two halves) in program
            // order.  The join barrier makes every error word readable.
            const int nl = S.n_lanes;
            const int nq = (nl == 4 && c.G >= 8 && c.args->lane_split == 4) ? 4 : 2;
            int j = 0;
            while (j + 1 < nq && c.rank >= (j + 1) * c.G / nq) ++j;
            const int r0 = j * c.G / nq, r1 = (j + 1) * c.G / nq;
            const unsigned mask = (nq == nl) ? (1u << j) : ((1u << j) | (1u << (j + 2)));
            if (nq == 2) {
                Ctx s = sub_team(c, r0, r1 - r0, j ? c.bar_hi : c.bar_lo, j ? c.tred_hi : c.tred_lo,
                                 j ? c.tpar_hi : c.tpar_lo);
                if (s.rank == 0) s.prof = S.prof;  // profiling: each sub-team's ops are timed by its rank 0
                s.err = S.error + 1 + j;
                run_lane(s, S, k, op.lane_end, mask, mode);
                (j ? c.tpar_hi : c.tpar_lo) = s.tpar;
            } else {
                Ctx s = sub_team(c, r0, r1 - r0, c.bar_q + j, c.tred_q + 2 * MAX_TEAM * j, c.tpar_q);
                if (s.rank == 0) s.prof = S.prof;
                s.err = S.error + 1 + j;
                run_lane(s, S, k, op.lane_end, mask, mode);
                c.tpar_q = s.tpar;
            }
            team_sync(c);
            int lane_err = 0;
            for (int q = 1; q <= nq; ++q) lane_err |= c.I[S.error + q];
            if (lane_err) {
                if (c.gtid == 0) c.I[c.err] = 1;
                team_sync(c);
                return false;
            }
            if (op.rot_graph_op >= 0) c.rot_graph_done = 1;
            k = op.lane_end;
            continue;
        }
        if (c.G > 1 && op.solo) {
            const int k2 = solo_run_end(S, k, k1);
            if (c.rank == 0) {
                Ctx l = solo_ctx(c);
                for (int q = k; q < k2; ++q) {
                    unsigned long long t0 = ptick(l);
                    run_value(l, S, S.ops[q], mode);
                    prof_op(l, q, t0);
                    if (l.I[l.err]) break;
                }
            }
            team_sync(c);
            if (c.I[c.err]) return false;
            k = k2;
            continue;
        }
        if (SPLIT && op.kind == OP_ROTAMER && op.overlap_end > k && op.overlap_end <= k1) {
            // A rotamer op with an overlap range: the ops of the range and the
            // early deriv ops run beside the solve where the team can split,
            // else between the solve and the rotamer's sensitivity push.
            // Either way they come after the solve's start and before the push,
            // so every sum sees its terms in the same order: same bits on
            // every team size and layout.
            unsigned long long t0 = ptick(c);
            const PairGraph& G = S.graphs[op.graph];
            const RotamerDev& R = S.rots[op.rot];
            if (!rotamer_pre(c, S, op, G, R)) return false;
            int it = 0;
            const bool one_wg = bp_on_one_wg(c, R, c.I[R.stats + 4], c.bp_layout);
            unsigned long long t1 = ptick(c);
            if (c.G > 1 && !one_wg && c.bp_layout != 1 && c.G >= 4) {
                // the state does not fit one workgroup's LDS: the first half
                // solves while the second half runs the range
                const int gb = c.G / 2;
                if (c.rank < gb) {
                    Ctx b = sub_team(c, 0, gb, c.bar_lo, c.tred_lo, c.tpar_lo);
                    it = rot_solve_team(b, R);
                    c.tpar_lo = b.tpar;
                } else {
                    Ctx o = sub_team(c, gb, c.G - gb, c.bar_hi, c.tred_hi, c.tpar_hi);
                    if (o.rank == 0) o.prof = S.prof;  // profiling: the ops of this half are timed by its rank 0
                    const unsigned long long t2 = S.prof ? wall_clock64() : 0ull;
                    if (run_germ_range<false>(o, S, k + 1, op.overlap_end, mode)) run_early_derivs(o, S);
                    c.tpar_hi = o.tpar;
                    // profiling: the half's own time, by its first thread (rank 0 keeps the others)
                    if (S.prof && o.gtid == 0)
                        __hip_atomic_fetch_add(S.prof + c.pbase + PROF_BESIDE, wall_clock64() - t2, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
                team_sync(c);  // join; an error in the range is read after the rotamer
            } else if (c.G > 1 && one_wg) {
                // rank 0 solves in its LDS while ranks 1..G-1 run the range
                if (c.rank == 0) {
                    it = rot_solve_wg(c, R);
                } else {
                    Ctx s = sub_team(c);
                    if (s.rank == 0) s.prof = S.prof;
                    const unsigned long long t2 = S.prof ? wall_clock64() : 0ull;
                    if (run_germ_range<false>(s, S, k + 1, op.overlap_end, mode)) run_early_derivs(s, S);
                    c.tpar2 = s.tpar;
                    if (S.prof && s.gtid == 0)
                        __hip_atomic_fetch_add(S.prof + c.pbase + PROF_BESIDE, wall_clock64() - t2, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
                team_sync(c);  // join; an error in the range is read after the rotamer
            } else {
                // one workgroup, or the whole team solves: the range follows
                it = rot_solve(c, R, c.bp_layout);
                if (!run_germ_range<false>(c, S, k + 1, op.overlap_end, mode)) return false;
                run_early_derivs(c, S);
            }
            pacc(c, PROF_BP, t1);
            c.early_done = 1;
            rotamer_post(c, S, op, G, R, mode, it);
            prof_op(c, k, t0);
            if (c.I[c.err]) return false;
            k = op.overlap_end;
            continue;
        }
        unsigned long long t0 = ptick(c);
        run_value(c, S, op, mode);
        prof_op(c, k, t0);
        if (c.I[c.err]) return false;  // uniform: read after the op's final barrier
        ++k;
    }
    return true;
}

// One evaluation: value ops in germ order, then sensitivity pushes.
__device__ void compute(Ctx& c, const SysDesc& S, int mode) {
    c.cur_op = S.n_germ + S.n_deriv;  // outside the ops
    c.early_done = 0;
    c.rot_graph_done = 0;
    const int* zr = c.I + S.sens_zero;
    for (int r = 0; r < S.sens_zero_cnt; ++r) {
        float* p = c.A + zr[2 * r];
        int n = zr[2 * r + 1];
        for (int i = c.gtid; i < n; i += c.gsize) p[i] = 0.f;
    }
    team_sync(c);
    if (!run_germ_range<true>(c, S, 0, S.n_germ, mode)) return;
    const int nd = S.n_germ + S.n_deriv;
    for (int k = S.n_germ; k < nd;) {
        c.cur_op = k;
        if (c.early_done && S.ops[k].early) { ++k; continue; }  // ran beside the rotamer solve
        if (c.G > 1 && S.ops[k].solo) {
            const int k2 = solo_run_end(S, k, nd);
            if (c.rank == 0) {
                Ctx l = solo_ctx(c);
                for (int q = k; q < k2; ++q) {
                    if (c.early_done && S.ops[q].early) continue;
                    unsigned long long t0 = ptick(l);
                    run_deriv(l, S, S.ops[q]);
                    prof_op(l, q, t0);
                }
            }
            team_sync(c);
            k = k2;
            continue;
        }
        unsigned long long t0 = ptick(c);
        run_deriv(c, S, S.ops[k]);
        prof_op(c, k, t0);
        ++k;
    }
}

// Threefry-4x32-20 (Salmon et al., SC'11), key (seed, stream, 0, 0) and counter
// (step lo, step hi, atom, 0) as RandomGenerator (src/random.h:19-67).
__device__ __forceinline__ uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }
__device__ void threefry4x32_20(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4]) {
    const int R0[8] = {10, 11, 13, 23, 6, 17, 25, 18};
    const int R1[8] = {26, 21, 27, 5, 20, 11, 10, 20};
    uint32_t ks[5] = {key[0], key[1], key[2], key[3], 0x1BD11BDAu ^ key[0] ^ key[1] ^ key[2] ^ key[3]};
    uint32_t x0 = ctr[0] + ks[0], x1 = ctr[1] + ks[1], x2 = ctr[2] + ks[2], x3 = ctr[3] + ks[3];
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        if ((r & 1) == 0) {
            x0 += x1; x1 = rotl32(x1, R0[r & 7]) ^ x0;
            x2 += x3; x3 = rotl32(x3, R1[r & 7]) ^ x2;
        } else {
            x0 += x3; x3 = rotl32(x3, R0[r & 7]) ^ x0;
            x2 += x1; x1 = rotl32(x1, R1[r & 7]) ^ x2;
        }
        if ((r & 3) == 3) {
            int s = (r >> 2) + 1;
            x0 += ks[s % 5]; x1 += ks[(s + 1) % 5]; x2 += ks[(s + 2) % 5]; x3 += ks[(s + 3) % 5] + (uint32_t)s;
        }
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

// normal3 of RandomGenerator: Box-Muller on u01/uneg11 (src/boxmuller.hpp, src/uniform.hpp).
__device__ void normal3(uint32_t seed, uint32_t stream, uint32_t atom, uint64_t step, float* o) {
    uint32_t key[4] = {seed, stream, 0u, 0u};
    uint32_t ctr[4] = {(uint32_t)(step & 0xffffffffull), (uint32_t)(step >> 32), atom, 0u};
    uint32_t b[4];
    threefry4x32_20(ctr, key, b);
    float a0 = (float)(int32_t)b[0] * 4.656612873077392578125e-10f + 2.3283064365386962890625e-10f;
    float u1 = (float)b[1] * 2.3283064365386962890625e-10f + 1.16415321826934814453125e-10f;
    float a2 = (float)(int32_t)b[2] * 4.656612873077392578125e-10f + 2.3283064365386962890625e-10f;
    float u3 = (float)b[3] * 2.3283064365386962890625e-10f + 1.16415321826934814453125e-10f;
    const float PIf = 3.1415926535897932f;
    float r1 = sqrtf(-2.f * logf(u1));
    float r3 = sqrtf(-2.f * logf(u3));
    float t1 = PIf * a0, t3 = PIf * a2;
    o[0] = sinf(t1) * r1;
    o[1] = cosf(t1) * r1;
    o[2] = sinf(t3) * r3;
}

// Temperature of the annealing schedule at `time` (the anneal_temp lambda of
// src/main.cpp:603-612, in double): T0 until the start, T0*factor after the
// end, between them the square of a linear blend of the square roots.
__device__ double anneal_temperature(const RunArgs& a, double T0, double time) {
    const double T1 = T0 * a.anneal_factor;
    if (time <= a.anneal_start) return T0;
    if (time > a.anneal_end) return T1;
    double f = (time - a.anneal_start) / a.anneal_duration;
    f = f > 0.0 ? f : 0.0;
    const double s = sqrt(T0) * (1.0 - f) + sqrt(T1) * f;
    return s * s;
}

__device__ void thermostat(Ctx& c, SysDesc& S, float noise_scale) {
    auto mom = gmem(c.A + S.mom);
    for (int a = c.gtid; a < S.n_atom; a += c.gsize) {
        float n[3];
        normal3(S.seed, 0u, (uint32_t)a, S.n_invocations, n);
        for (int k = 0; k < 3; ++k) mom[a * 3 + k] = S.mom_scale * mom[a * 3 + k] + noise_scale * n[k];
    }
    team_sync(c);
    if (c.gtid == 0) S.n_invocations += 1;
    team_sync(c);
}

__device__ float total_potential(Ctx& c, const SysDesc& S) {
    float t = 0.f;
    for (int k = 0; k < S.n_pot; ++k) t += c.A[S.pot + k];
    return t;
}

// Barrier over every workgroup of the launch that takes part in replica
// exchange (teams on different XCDs: device-scope fences on both sides), on a
// monotonic arrival counter as team_sync.  Bounded: a timeout sets the flag,
// which the host turns into an error, and no later wait blocks.
__device__ void grid_sync(Ctx& c, const RexDev& x) {
    __syncthreads();
    if (c.tid == 0) {
        __threadfence();
        unsigned long long old = __hip_atomic_fetch_add(x.grid_bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long target = (old / (unsigned long long)x.n_wg + 1ull) * (unsigned long long)x.n_wg;
        unsigned spins = 0;
        while (__hip_atomic_load(x.grid_bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > BAR_SPIN_LIMIT ||
                __hip_atomic_load(x.grid_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                __hip_atomic_store(x.grid_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// One replica-exchange attempt after `done` completed rounds (see RexDev):
// all teams of the launch.  The decisions follow attempt_swaps
// (src/main.cpp:243-316): criterion 0 keeps a swap unless the change of the
// summed log Boltzmann weights is negative and exp(change) falls below a
// uniform draw; criterion 1 swaps when the change lb[s2] - lb[s1] is
// non-negative or exp(change) reaches the draw; draws (0, 1] from Threefry
// stream 1, counter (done, 0, k), one per test that needs it.  With one
// shared potential the weights after a swap are those of the partner's
// positions, -beta_s E(x_partner), evaluated once per system.
__device__ __noinline__ void rex_attempt(Ctx& c, SysDesc& S, const RexDev& x, int sys, int n_sys, long long done) {
    compute(c, S, 1);
    if (c.gtid == 0) x.energy[sys] = total_potential(c, S);
    {
        auto pos = gmem(c.A + S.pos_out);
        auto st = gmem(x.stage + (size_t)sys * x.stage_stride);
        for (int i = c.gtid; i < S.n_atom * 3; i += c.gsize) st[i] = pos[i];
    }
    grid_sync(c, x);
    if (sys == 0 && c.rank == 0 && c.tid == 0) {
        const long long att = x.state[0];
        const int set = (int)(att % (long long)x.n_sets);
        for (int k = 0; k < n_sys; ++k) x.partner[k] = -1;
        uint32_t key[4] = {x.seed, 1u, 0u, 0u};
        uint32_t ctr[4] = {(uint32_t)((unsigned long long)done & 0xffffffffull), (uint32_t)((unsigned long long)done >> 32), 0u, 0u};
        for (int p = x.set_off[set]; p < x.set_off[set + 1]; ++p) {
            const int s1 = x.pairs[2 * p], s2 = x.pairs[2 * p + 1];
            const float e1 = x.energy[s1], e2 = x.energy[s2];
            const float b1 = x.beta[s1], b2 = x.beta[s2];
            bool swap;
            auto uniform = [&]() {
                uint32_t bits[4];
                threefry4x32_20(ctr, key, bits);
                ctr[3]++;
                return (float)bits[0] * 2.3283064365386962890625e-10f + 1.16415321826934814453125e-10f;
            };
            if (x.criterion == 0) {
                const float diff = ((-b1 * e2) + (-b2 * e1)) - ((-b1 * e1) + (-b2 * e2));
                swap = !(diff < 0.f && expf(diff) < uniform());
            } else {
                const float diff = (-b2 * e2) - (-b1 * e1);
                swap = diff >= 0.f || expf(diff) >= uniform();
            }
            x.stats[2 * p + 1] += 1;
            if (swap) {
                x.stats[2 * p] += 1;
                x.partner[s1] = s2;
                x.partner[s2] = s1;
                const int t = x.replica_of[s1];
                x.replica_of[s1] = x.replica_of[s2];
                x.replica_of[s2] = t;
            }
        }
        x.state[0] = att + 1;
        // the next attempt ends the next segment of the reference's loop: the
        // first multiple of the interval at least two rounds on (:918-990)
        const long long e = done + 2;
        x.state[1] = (e + x.interval - 1) / x.interval * x.interval;
    }
    grid_sync(c, x);
    const int q = x.partner[sys];
    if (q >= 0) {
        auto pos = gmem(c.A + S.pos_out);
        auto st = gmem((const float*)x.stage + (size_t)q * x.stage_stride);
        for (int i = c.gtid; i < S.n_atom * 3; i += c.gsize) pos[i] = st[i];
    }
    team_sync(c);
}

// One launch evaluates or integrates every system of a batch; system s is
// handled by the team of workgroups [s*team, (s+1)*team).  The body is shared
// by two kernel symbols so that the integration launches (the hot path) and
// the single evaluations / thermostat applications show up separately in a
// kernel trace.
__device__ __forceinline__ void engine_body(SysDesc* systems, const RunArgs& args, float* red, int* ired,
                                            float* lds) {
    int G = args.team;
    int sys = blockIdx.x, rank = 0;
    if (args.blockmap) {  // per-system teams, placed by the host
        const int e = args.blockmap[blockIdx.x];
        if (e < 0) return;  // padding workgroup
        sys = e >> 14;
        G = (e >> 7) & 127;
        rank = e & 127;
    } else if (G > 1) {  // team launches: workgroup b runs on XCD b % N_XCD; a team stays on one XCD
        const int xcd = blockIdx.x % N_XCD, slot = blockIdx.x / N_XCD;
        sys = (slot / G) * N_XCD + xcd;
        rank = slot % G;
        if (sys >= args.n_sys) return;  // padding workgroup
    }
    SysDesc& S = systems[sys];
    Ctx c;
    c.A = S.arena; c.I = S.iarena; c.CF = S.cf; c.CI = S.ci;
    c.red = red; c.ired = ired; c.lds = lds;
    c.tid = (int)threadIdx.x;
    c.prof = rank == 0 ? S.prof : nullptr;
    c.pbase = S.n_germ + S.n_deriv;
    c.cur_op = c.pbase;
    c.err = S.error;
    c.rank = rank; c.G = G;
    c.gtid = rank * NT + c.tid; c.gsize = G * NT;
    c.bar = reinterpret_cast<unsigned long long*>(S.iarena + S.bar);  // 16-byte aligned words
    c.bar_flag = reinterpret_cast<unsigned*>(S.iarena + S.bar + 2);
    c.bar_next = (G > 1 && c.tid == 0) ? bar_first(c.bar, G) : 0ull;
    c.tred = S.arena + S.tred;
    c.tpar = 0;
    c.bar2 = reinterpret_cast<unsigned long long*>(S.iarena + S.bar + BAR_SUB1);
    c.tred2 = c.tred + 2 * MAX_TEAM;
    c.tpar2 = 0;
    c.bar_lo = reinterpret_cast<unsigned long long*>(S.iarena + S.bar + BAR_LO);
    c.tred_lo = c.tred + 4 * MAX_TEAM;
    c.tpar_lo = 0;
    c.bar_hi = reinterpret_cast<unsigned long long*>(S.iarena + S.bar + BAR_HI);
    c.tred_hi = c.tred + 6 * MAX_TEAM;
    c.tpar_hi = 0;
    c.bar_q = reinterpret_cast<unsigned long long*>(S.iarena + S.bar + BAR_Q);
    c.tred_q = c.tred + 8 * MAX_TEAM;
    c.tpar_q = 0;
    c.bp_layout = args.bp_layout;
    c.args = &args;
    team_place(c, S.iarena + S.bar + 4);
    if (args.n_round == 0) {
        compute(c, S, args.mode);
        return;
    }
    if (args.n_round < 0) {  // thermostat application only (fresh-start thermalisation)
        thermostat(c, S, S.noise_scale);
        return;
    }
    auto pos = gmem(c.A + S.pos_out);
    auto der = gmem(c.A + S.pos_sens);
    auto mom = gmem(c.A + S.mom);
    int nf = 0;
    for (int rr = 0; rr < args.n_round; ++rr) {
        int round = args.first_round + rr;
        if (args.frame_interval > 0 && round % args.frame_interval == 0) {
            compute(c, S, 1);
            if (c.I[c.err]) return;
            if (args.frame_pot && c.gtid == 0 && nf < args.n_frames)
                args.frame_pot[sys * args.n_frames + nf] = total_potential(c, S);
            ++nf;
            team_sync(c);
        }
        unsigned long long t0 = ptick(c);
        if (args.thermostat_interval > 0 && round % args.thermostat_interval == 0) {
            float ns = S.noise_scale;
            if (args.anneal) {  // set_temperature then the update_parameters of src/thermostat.h:9-12
                const float time = 3.f * args.dt * (float)(round + 1);
                const float temp = (float)anneal_temperature(args, (double)S.temp0, (double)time);
                ns = sqrtf(temp * (1.f - S.mom_scale * S.mom_scale));
            }
            thermostat(c, S, ns);
        }
        pacc(c, PROF_THERMO, t0);
        for (int stage = 0; stage < 3; ++stage) {
            compute(c, S, 0);
            if (c.I[c.err]) return;  // capacity overflow: the host repeats the run
            t0 = ptick(c);
            c.cur_op = S.n_germ + S.n_deriv;
            for (int i = c.gtid; i < S.n_atom * 3; i += c.gsize) {
                float p = mom[i] - args.dt * der[i];
                mom[i] = p;
                pos[i] += args.dt * p;
            }
            team_sync(c);
            pacc(c, PROF_INTEGRATE, t0);
        }
        // replica exchange where a segment of the reference's loop ends
        // (src/main.cpp:991-997): after this round the count of completed
        // rounds is round + 1
        if (args.rex.interval > 0 && (long long)round + 1 == args.rex.state[1]) {
            rex_attempt(c, S, args.rex, sys, args.n_sys, (long long)round + 1);
            if (c.I[c.err]) return;
        }
    }
}

// Integration rounds (args.n_round > 0).
__global__ void __launch_bounds__(NT) engine_step_kernel(SysDesc* systems, RunArgs args) {
    __shared__ float red[NWAVE];
    __shared__ int ired[NWAVE];
    extern __shared__ float lds[];
    engine_body(systems, args, red, ired, lds);
}

// One evaluation (args.n_round == 0) or one thermostat application (< 0).
__global__ void __launch_bounds__(NT) engine_eval_kernel(SysDesc* systems, RunArgs args) {
    __shared__ float red[NWAVE];
    __shared__ int ired[NWAVE];
    extern __shared__ float lds[];
    engine_body(systems, args, red, ired, lds);
}

// ---- parameter derivatives of spline pair graphs ------------------------------
// Weights of the four coefficients of a uniform cubic B-spline at x >= 1 and
// the first coefficient they multiply: the derivative of the value with
// respect to each coefficient (deBoor_coeff_deriv, src/spline.h:366-384).
__device__ __forceinline__ int bspline_weights(float x, float* w) {
    const int xb = (int)x;
    const float y = x - (float)xb;
    const float omy = 1.f - y, y2 = y * y, y3 = y2 * y;
    w[0] = omy * omy * omy * (1.f / 6.f);
    w[1] = (3.f * y3 - 6.f * y2 + 4.f) * (1.f / 6.f);
    w[2] = (-3.f * y3 + 3.f * y2 + 3.f * y + 1.f) * (1.f / 6.f);
    w[3] = y3 * (1.f / 6.f);
    return xb - 1;
}

// Clamped variant, constant outside [1, n_knot-2] (clamped_deBoor_coeff_deriv,
// src/spline.h:423-440); NaN takes the lower end.
__device__ __forceinline__ int clamped_bspline_weights(float x, int n_knot, float* w) {
    if (!(x > 1.f)) {
        w[0] = 1.f / 6.f; w[1] = 2.f / 3.f; w[2] = 1.f / 6.f; w[3] = 0.f;
        return 0;
    }
    if (x >= (float)(n_knot - 2)) {
        w[0] = 0.f; w[1] = 1.f / 6.f; w[2] = 2.f / 3.f; w[3] = 1.f / 6.f;
        return n_knot - 4;
    }
    return bspline_weights(x, w);
}

// scale * d(quadspline)/d(p) added into acc[0..2*NKA+2*NK): the wide radial
// spline, the narrow one times both angular values, each angular spline times
// the other and the narrow value (quadspline_param_deriv,
// src/bead_interaction.h:87-131).  Geometry as in the value (quadspline).
template <int NKA, int NK>
__device__ void quadspline_param_deriv(const float* p, const float* x1, const float* x2, float inv_dtheta,
                                       float inv_dx, float scale, float* acc) {
    float dx = x2[0] - x1[0], dy = x2[1] - x1[1], dz = x2[2] - x1[2];
    float dist2 = __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, dx * dx));
    float inv = rsqrtf(dist2);
    float dist_coord = dist2 * (inv * inv_dx);
    float ux = inv * dx, uy = inv * dy, uz = inv * dz;
    float c1 = x1[3] * ux + x1[4] * uy + x1[5] * uz;
    float c2 = -(x2[3] * ux + x2[4] * uy + x2[5] * uz);
    float xa1 = fminf(fmaxf((c1 + 1.f) * inv_dtheta + 1.f, 1.f), (float)(NKA - 2));
    float xa2 = fminf(fmaxf((c2 + 1.f) * inv_dtheta + 1.f, 1.f), (float)(NKA - 2));
    float a1, a2, nn, unused;
    deboor(p, xa1, a1, unused);
    deboor(p + NKA, xa2, a2, unused);
    clamped_deboor(p + 2 * NKA + NK, dist_coord, NK, nn, unused);
    float w[4];
    int b = clamped_bspline_weights(dist_coord, NK, w);
    for (int i = 0; i < 4; ++i) acc[2 * NKA + b + i] += scale * w[i];
    for (int i = 0; i < 4; ++i) acc[2 * NKA + NK + b + i] += scale * (a1 * a2 * w[i]);
    b = bspline_weights(xa1, w);
    for (int i = 0; i < 4; ++i) acc[b + i] += scale * (a2 * nn * w[i]);
    b = bspline_weights(xa2, w);
    for (int i = 0; i < 4; ++i) acc[NKA + b + i] += scale * (a1 * nn * w[i]);
}

// One workgroup per interaction-table row t: every thread walks the slots in
// a fixed stride, adds the derivative of each present edge whose type pair is
// t (or, in a symmetric graph, whose transposed pair is t) into its own LDS
// row, and the rows are summed in thread order, so the result is the same on
// every run.  In a symmetric graph an edge is taken with its lower element
// first, the orientation the reference's pair list gives it
// (src/interaction_graph.h:134-145).
__global__ void __launch_bounds__(NT) param_deriv_kernel(ParamDerivArgs a) {
    extern __shared__ float lds[];
    const int t = blockIdx.x;
    const int np = a.n_param;
    float* acc = lds + threadIdx.x * np;
    for (int d = 0; d < np; ++d) acc[d] = 0.f;
    for (int s = threadIdx.x; s < a.n_slot; s += NT) {
        if (!a.act[s]) continue;
        int i = a.row[s], j = a.nbr[s];
        if (a.sym && j < i) { int k = i; i = j; j = k; }
        const int t1 = a.type1[i], t2 = a.type2[j];
        const int m = (t1 * a.n_type2 + t2 == t) + (a.sym && t2 * a.n_type2 + t1 == t);
        if (!m) continue;
        const float* x1 = a.x1 + a.loc1[i] * a.w1;
        const float* x2 = a.x2 + a.loc2[j] * a.w2;
        const float* p = a.param + (t1 * a.n_type2 + t2) * np;
        for (int k = 0; k < m; ++k) {
            if (a.kind == EDGE_ROTAMER) {
                quadspline_param_deriv<15, 12>(p, x1, x2, 6.f, 1.f / 0.7f, a.es[s], acc);
            } else {  // hbond coverage: (1 - hb)^2 times the spline (src/hbond.cpp:286-291)
                const float om = 1.f - x1[6];
                quadspline_param_deriv<15, 10>(p, x1, x2, 6.f, 1.f / 0.625f, a.es[s] * (om * om), acc);
            }
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < np; d += NT) {
        float sum = 0.f;
        for (int k = 0; k < NT; ++k) sum += lds[k * np + d];
        a.out[t * np + d] = sum;
    }
}

// Sigmoid coupling (SigmoidCoupling::get_param_deriv,
// src/environment.cpp:958-997): entries [0,nt) d/dscale, [nt,2nt) d/dcenter,
// [2nt,3nt) d/dsharpness per coupling type, then the weight block when the
// node has independent weights (400: per coupling type and amino acid; 20: per
// amino acid, repeated).  Backbone sigmoid coupling
// (BackboneSigmoidCoupling::get_param_deriv, :1118-1134): only d/dscale.
// Each thread owns output entries and sums over residues in residue order.
__global__ void __launch_bounds__(NT) coupling_param_deriv_kernel(CouplingDerivArgs a) {
    const int nt = a.nt;
    for (int j = threadIdx.x; j < a.n_out; j += NT) {
        float acc = 0.f;
        if (a.kind == OP_HBOND_ENERGY) {
            // HBondEnergy::get_param_deriv (src/hbond.cpp:655-720): [0..2] hbonds of
            // each residue times its helix, sheet and turn scores, [3] the bias
            // term of each peptide; the shape parameters [4..11] stay zero
            const float* P = a.hp;
            if (j < 3) {
                for (int i = 0; i < a.n_res; ++i) {
                    const float phi = a.rama[i * a.wr], psi = a.rama[i * a.wr + 1];
                    float p1, d1, p2, d2, s1, e1, s2, e2;
                    compact_sigmoid(phi - P[4], P[5], p1, d1);
                    compact_sigmoid(phi - P[6], P[7], p2, d2);
                    compact_sigmoid(psi - P[8], P[9], s1, e1);
                    compact_sigmoid(psi - P[10], P[11], s2, e2);
                    const float turn = (float)((1. - (double)p1) * p2);
                    const float psis = (float)((1. - (double)s1) * s2);
                    const float nturn = (float)(1. - (double)turn), npsi = (float)(1. - (double)psis);
                    const float score = j == 0 ? nturn * psis : (j == 1 ? nturn * npsi : turn);
                    float hb = 0.f;
                    for (int k = a.h1o[i]; k < a.h1o[i + 1]; ++k) hb += a.pp[a.h1i[k] * a.wp + 6];
                    acc += hb * score;
                }
            } else if (j == 3) {
                for (int i = 0; i + 1 < a.n_res; ++i) {
                    float hb = 0.f;
                    for (int k = a.h2o[i]; k < a.h2o[i + 1]; ++k) hb += a.pp[a.h2i[k] * a.wp + 6];
                    float v, d;
                    compact_sigmoid(1.5f - hb, 2.0f, v, d);
                    acc += v;
                }
            }
        } else if (a.kind == OP_BB_SIGMOID_COUPLING) {
            if (j == 0)
                for (int r = 0; r < a.n_res; ++r) {
                    float bl = 0.f;
                    for (int aa = 0; aa < nt; ++aa) bl += a.weights[aa] * a.cov[r * nt + aa];
                    bl += a.bb_hbond_weight * a.cov_hb[r];
                    float v, d;
                    compact_sigmoid(bl - a.bb_center, a.bb_sharp, v, d);
                    acc += v;
                }
        } else {
            const int cs = 3 * nt;
            const int q = j < cs ? j / nt : 3;
            const int want_ct = j < cs ? j % nt : (a.n_indep == 400 ? (j - cs) / nt : -1);
            const int want_aa = j < cs ? -1 : (j - cs) % nt;
            if (j < cs || a.n_indep > 1)
                for (int r = 0; r < a.n_res; ++r) {
                    const int ct = a.ctype[r];
                    if (want_ct >= 0 && ct != want_ct) continue;
                    float w = 0.f;
                    for (int aa = 0; aa < nt; ++aa) w += a.weights[ct * nt + aa] * a.cov[r * nt + aa];
                    const float dc = w - a.center[ct];
                    float v, d;
                    compact_sigmoid(dc, a.sharp[ct], v, d);
                    if (q == 0) acc += v;
                    else if (q == 1) acc += -a.scale[ct] * d;
                    else if (q == 2) acc += a.scale[ct] * dc * d / a.sharp[ct];
                    else acc += (d * a.scale[ct]) * a.cov[r * nt + want_aa];
                }
        }
        a.out[j] = acc;
    }
}

// Subtract the centroid of the positions (x, y and z, or x and y only), as
// recenter does before a logged frame (src/deriv_engine.cpp:39-49): the sum
// runs over atoms in order on one thread, so it rounds as the reference's.
__global__ void __launch_bounds__(NT) recenter_kernel(float* pos, int n_atom, int xy_only) {
    __shared__ float cen[3];
    if (threadIdx.x == 0) {
        float cx = 0.f, cy = 0.f, cz = 0.f;
        for (int a = 0; a < n_atom; ++a) { cx += pos[3 * a]; cy += pos[3 * a + 1]; cz += pos[3 * a + 2]; }
        const float inv = (float)n_atom;
        cen[0] = cx / inv;
        cen[1] = cy / inv;
        cen[2] = xy_only ? 0.f : cz / inv;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * n_atom; i += NT) pos[i] -= cen[i % 3];
}

}  // namespace upside

hipError_t upside_launch_recenter(float* pos, int n_atom, int xy_only, hipStream_t stream) {
    if (n_atom <= 0) return hipErrorInvalidValue;
    void* kargs[] = {&pos, &n_atom, &xy_only};
    return hipLaunchKernel((const void*)upside::recenter_kernel, dim3(1), dim3(upside::NT), kargs, 0, stream);
}

hipError_t upside_launch_coupling_param_deriv(const upside::CouplingDerivArgs& args, hipStream_t stream) {
    if (args.n_out <= 0 || args.nt <= 0) return hipErrorInvalidValue;
    upside::CouplingDerivArgs a = args;
    void* kargs[] = {&a};
    return hipLaunchKernel((const void*)upside::coupling_param_deriv_kernel, dim3(1), dim3(upside::NT), kargs, 0,
                           stream);
}

hipError_t upside_launch_param_deriv(const upside::ParamDerivArgs& args, int n_rows, hipStream_t stream) {
    const size_t lds_bytes = (size_t)upside::NT * args.n_param * sizeof(float);
    if (args.n_param <= 0 || lds_bytes > upside::LDS_FLOATS * sizeof(float) || n_rows <= 0) return hipErrorInvalidValue;
    (void)hipFuncSetAttribute((const void*)upside::param_deriv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
    upside::ParamDerivArgs a = args;
    void* kargs[] = {&a};
    return hipLaunchKernel((const void*)upside::param_deriv_kernel, dim3(n_rows), dim3(upside::NT), kargs, lds_bytes,
                           stream);
}

// Host-side launcher (C++ linkage inside the library only).
hipError_t upside_launch_engine(upside::SysDesc* d_systems, int n_sys, upside::RunArgs args, hipStream_t stream) {
    static bool attr_set = false;
    const size_t lds_bytes = upside::LDS_FLOATS * sizeof(float);
    if (!attr_set) {
        for (const void* f : {(const void*)upside::engine_step_kernel, (const void*)upside::engine_eval_kernel}) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
            if (e != hipSuccess) return e;
        }
        attr_set = true;
    }
    const void* fn = args.n_round > 0 ? (const void*)upside::engine_step_kernel
                                      : (const void*)upside::engine_eval_kernel;
    if (args.team <= 1) args.team = 1;
    args.n_sys = n_sys;
    void* kargs[] = {&d_systems, &args};
    if (args.blockmap)  // per-system teams: every workgroup resident at once
        return hipLaunchCooperativeKernel(fn, dim3(args.n_blocks), dim3(upside::NT), kargs, (unsigned)lds_bytes,
                                          stream);
    if (args.team == 1) {
        if (args.rex.interval > 0)  // the grid barriers need every workgroup resident
            return hipLaunchCooperativeKernel(fn, dim3(n_sys), dim3(upside::NT), kargs, (unsigned)lds_bytes, stream);
        return hipLaunchKernel(fn, dim3(n_sys), dim3(upside::NT), kargs, lds_bytes, stream);
    }
    // teams wait for each other: every workgroup must be resident at once,
    // which a cooperative launch guarantees (or refuses).  The grid covers
    // N_XCD slots per team rank; slots past the last system exit at once.
    const int groups = (n_sys + upside::N_XCD - 1) / upside::N_XCD;
    return hipLaunchCooperativeKernel(fn, dim3(upside::N_XCD * args.team * groups), dim3(upside::NT), kargs,
                                      (unsigned)lds_bytes, stream);
}

// Workgroups of the engine kernels the device can hold at once (the smaller
// of the two symbols' occupancies).
int upside_max_resident_blocks() {
    static int cached = -1;
    if (cached >= 0) return cached;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    const size_t lds_bytes = upside::LDS_FLOATS * sizeof(float);
    int best = 1 << 30;
    for (const void* f : {(const void*)upside::engine_step_kernel, (const void*)upside::engine_eval_kernel}) {
        int per_cu = 0;
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, upside::NT, lds_bytes) != hipSuccess) return 1;
        best = per_cu < best ? per_cu : best;
    }
    cached = best * cus;
    return cached;
}
