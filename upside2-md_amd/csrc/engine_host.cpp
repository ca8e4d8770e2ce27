// Host side of the MI355X Upside engine: `.up` loading, lowering of the node
// graph to a device program, device memory, and the C ABI of include/upside_engine.h.
//
// Loading follows initialize_engine_from_hdf5 (src/deriv_engine.cpp:353-426):
// nodes under /input/potential are ordered topologically by their `arguments`
// attribute, typed by registered name prefix (:393-399), then germ and deriv
// execution levels are built as build_exec_levels does (:127-178).  Every node
// type of the ff_2.1 graph is lowered; any other type makes construction fail
// with a message, as an unknown type does in the reference.
#include <hdf5.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_dev.h"
#include "warmstart.h"
#include "../../include/upside_engine.h"

hipError_t upside_launch_engine(upside::SysDesc* d_systems, int n_sys, upside::RunArgs args, hipStream_t stream);
hipError_t upside_launch_param_deriv(const upside::ParamDerivArgs& args, int n_rows, hipStream_t stream);
hipError_t upside_launch_coupling_param_deriv(const upside::CouplingDerivArgs& args, hipStream_t stream);
hipError_t upside_launch_recenter(float* pos, int n_atom, int xy_only, hipStream_t stream);
int upside_max_resident_blocks();

namespace upside {

using std::string;
using std::vector;

[[noreturn]] static void fail(const string& s) { throw std::runtime_error(s); }

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ HDF5 ---
struct H5 {
    hid_t f = -1;
    explicit H5(const string& path) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
        if (f < 0) fail("unable to open " + path);
    }
    ~H5() { if (f >= 0) H5Fclose(f); }
    bool exists(const string& p) const {
        string cur;
        size_t i = 1;
        while (i <= p.size()) {
            size_t j = p.find('/', i);
            if (j == string::npos) j = p.size();
            cur = p.substr(0, j);
            if (H5Lexists(f, cur.c_str(), H5P_DEFAULT) <= 0) return false;
            i = j + 1;
        }
        return true;
    }
    vector<size_t> shape(const string& p) const {
        hid_t d = H5Dopen2(f, p.c_str(), H5P_DEFAULT);
        if (d < 0) fail("missing dataset " + p);
        hid_t s = H5Dget_space(d);
        int nd = H5Sget_simple_extent_ndims(s);
        vector<hsize_t> dims(nd > 0 ? nd : 0);
        if (nd > 0) H5Sget_simple_extent_dims(s, dims.data(), nullptr);
        H5Sclose(s);
        H5Dclose(d);
        return vector<size_t>(dims.begin(), dims.end());
    }
    template <typename T>
    vector<T> read(const string& p, hid_t mt) const {
        auto sh = shape(p);
        size_t n = 1;
        for (auto x : sh) n *= x;
        vector<T> out(n);
        if (!n) return out;
        hid_t d = H5Dopen2(f, p.c_str(), H5P_DEFAULT);
        if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0) {
            H5Dclose(d);
            fail("unable to read " + p);
        }
        H5Dclose(d);
        return out;
    }
    vector<float> rf(const string& p) const { return read<float>(p, H5T_NATIVE_FLOAT); }
    vector<double> rd(const string& p) const { return read<double>(p, H5T_NATIVE_DOUBLE); }
    vector<int> ri(const string& p) const { return read<int>(p, H5T_NATIVE_INT); }
    bool has_attr(const string& p, const string& a) const {
        return H5Aexists_by_name(f, p.c_str(), a.c_str(), H5P_DEFAULT) > 0;
    }
    double attr_d(const string& p, const string& a) const {
        if (!has_attr(p, a)) fail("missing attribute " + a + " of " + p);
        double v = 0;
        hid_t at = H5Aopen_by_name(f, p.c_str(), a.c_str(), H5P_DEFAULT, H5P_DEFAULT);
        H5Aread(at, H5T_NATIVE_DOUBLE, &v);
        H5Aclose(at);
        return v;
    }
    float attr_f(const string& p, const string& a) const {
        float v = 0;
        if (!has_attr(p, a)) fail("missing attribute " + a + " of " + p);
        hid_t at = H5Aopen_by_name(f, p.c_str(), a.c_str(), H5P_DEFAULT, H5P_DEFAULT);
        H5Aread(at, H5T_NATIVE_FLOAT, &v);
        H5Aclose(at);
        return v;
    }
    int attr_i(const string& p, const string& a, int dflt, bool required) const {
        if (!has_attr(p, a)) {
            if (required) fail("missing attribute " + a + " of " + p);
            return dflt;
        }
        int v = 0;
        hid_t at = H5Aopen_by_name(f, p.c_str(), a.c_str(), H5P_DEFAULT, H5P_DEFAULT);
        H5Aread(at, H5T_NATIVE_INT, &v);
        H5Aclose(at);
        return v;
    }
    vector<string> attr_strings(const string& p, const string& a) const {
        if (!has_attr(p, a)) fail("missing attribute " + a + " of " + p);
        hid_t at = H5Aopen_by_name(f, p.c_str(), a.c_str(), H5P_DEFAULT, H5P_DEFAULT);
        hid_t t = H5Aget_type(at);
        hid_t s = H5Aget_space(at);
        if (H5Tis_variable_str(t) > 0) fail("variable-length strings not supported");
        size_t len = H5Tget_size(t);
        int nd = H5Sget_simple_extent_ndims(s);
        hsize_t n = 1;
        if (nd == 1) H5Sget_simple_extent_dims(s, &n, nullptr);
        vector<char> buf(n * len + 1, 0);
        if (n) H5Aread(at, t, buf.data());
        vector<string> out;
        for (hsize_t i = 0; i < n; ++i) {
            string x(buf.data() + i * len, len);
            while (!x.empty() && x.back() == '\0') x.pop_back();
            out.push_back(x);
        }
        H5Sclose(s);
        H5Tclose(t);
        H5Aclose(at);
        return out;
    }
    vector<string> children(const string& p) const {
        vector<string> out;
        hid_t g = H5Gopen2(f, p.c_str(), H5P_DEFAULT);
        if (g < 0) fail("missing group " + p);
        H5G_info_t info;
        H5Gget_info(g, &info);
        for (hsize_t i = 0; i < info.nlinks; ++i) {
            ssize_t n = H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
            vector<char> b(n + 1);
            H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, b.data(), n + 1, H5P_DEFAULT);
            out.emplace_back(b.data());
        }
        H5Gclose(g);
        return out;
    }
};

// --------------------------------------------------------------- layout ---
struct Node {
    string name, type;
    vector<string> args;
    vector<int> parents, children;
    bool potential = false;
    int n_elem = 0, width = 0;
    int out = -1, sens = -1;
    int pot_slot = -1;
    int op = -1;
    int germ_level = -1, deriv_level = -1;
};

struct Topology {
    vector<Node> nodes;
    vector<OpRec> ops;           // one per non-pos node, in node order
    vector<int> germ, deriv;     // node indices
    vector<PairGraph> graphs;
    vector<RotamerDev> rots;
    vector<int> graph_node;      // graph -> node
    vector<float> cf;
    vector<int> ci;
    int nf = 0, ni = 0;          // per-system arena sizes
    vector<std::pair<int, int>> sens_zero;
    vector<int> sens_zero_host;
    int n_pot = 0, pot = -1, mom = -1, err = -1, szero = -1, bar = -1, tred = -1;
    int n_atom = 0;
    double cap_scale = 1.0;

    int falloc(long n) { int o = nf; nf += (int)((n + 3) / 4 * 4); return o; }
    int ialloc(long n) { int o = ni; ni += (int)((n + 3) / 4 * 4); return o; }
    int cf_push(const vector<float>& v, int pad = 0) {
        int o = (int)cf.size() + pad;
        cf.insert(cf.end(), pad, 0.f);
        cf.insert(cf.end(), v.begin(), v.end());
        cf.insert(cf.end(), pad + 4, 0.f);
        return o;
    }
    int ci_push(const vector<int>& v) {
        int o = (int)ci.size();
        ci.insert(ci.end(), v.begin(), v.end());
        ci.insert(ci.end(), 4, 0);
        return o;
    }
    // Plan: target t of a producer receives the sum of contributions in order.
    int plan(int n_target, int cw, int tsens, int tstride, int col, int cbase, int cstride,
             const vector<std::pair<int, int>>& contrib) {
        vector<int> cnt(n_target + 1, 0);
        for (auto& p : contrib) {
            if (p.first < 0 || p.first >= n_target) fail("gather plan target out of range");
            cnt[p.first + 1]++;
        }
        for (int t = 0; t < n_target; ++t) cnt[t + 1] += cnt[t];
        vector<int> idx(contrib.size());
        vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (auto& p : contrib) idx[fill[p.first]++] = p.second;
        vector<int> h = {n_target, cw, tsens, tstride, col, cbase, cstride, (int)contrib.size()};
        h.insert(h.end(), cnt.begin(), cnt.end());
        h.insert(h.end(), idx.begin(), idx.end());
        return ci_push(h);
    }
};

static OpRec blank_op(int kind, int node) {
    OpRec o;
    std::memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.node = node;
    o.out = o.sens = o.pot_slot = -1;
    o.scratch = o.iscratch = -1;
    o.graph = o.rot = -1;
    for (int k = 0; k < 4; ++k) { o.in_out[k] = o.in_sens[k] = -1; o.plans[k] = -1; }
    return o;
}

static const char* kPrefixes[] = {
    "Distance3D", "Angle", "Dihedral", "Spring", "rama_coord", "cat_pos", "rama_map_pot",
    "affine_alignment", "placement_scalar", "placement_fixed_point_vector_only",
    "placement_fixed_point_vector_scalar", "infer_H_O", "protein_hbond", "hbond_coverage",
    "hbond_energy", "rotamer", "weighted_pos", "environment_coverage", "hb_environment_coverage",
    "sigmoid_coupling", "bb_sigmoid_coupling", "backbone_pairs"};

static string node_type(const string& name) {
    string best;
    for (auto p : kPrefixes) {
        string s(p);
        if (name.compare(0, s.size(), s) == 0 && s.size() > best.size()) best = s;
    }
    return best;
}

struct Builder {
    Topology& T;
    H5& h5;
    explicit Builder(Topology& t, H5& h) : T(t), h5(h) {}

    Node& nd(int i) { return T.nodes[i]; }
    string gp(const Node& n) const { return "/input/potential/" + n.name; }

    void alloc_coord(Node& n, int n_elem, int width) {
        n.n_elem = n_elem;
        n.width = width;
        n.out = T.falloc((long)n_elem * width);
        n.sens = T.falloc((long)n_elem * width);
        T.sens_zero.push_back({n.sens, n_elem * width});
    }
    void set_inputs(OpRec& o, const Node& n) {
        for (size_t k = 0; k < n.parents.size() && k < 4; ++k) {
            const Node& p = T.nodes[n.parents[k]];
            o.in_out[k] = p.out;
            o.in_sens[k] = p.sens;
            o.in_w[k] = p.width;
        }
    }
    const Node& parent(const Node& n, int k) { return T.nodes[n.parents[k]]; }
    void need_args(const Node& n, size_t k) {
        if (n.parents.size() != k) fail("node " + n.name + " expects " + std::to_string(k) + " arguments");
    }
    void check_range(const vector<int>& v, int lim, const string& what) {
        for (int x : v)
            if (x < 0 || x >= lim) fail(what + " index out of range");
    }
    // contributions (e*K+k) -> target rows of the producer for all slots k
    int atom_plan(const Node& prod, const vector<int>& ids, int K, int cbase, int cw) {
        vector<std::pair<int, int>> con;
        for (size_t r = 0; r < ids.size(); ++r)
            if (ids[r] >= 0) con.push_back({ids[r], (int)r});
        return T.plan(prod.n_elem, cw, prod.sens, prod.width, 0, cbase, cw, con);
    }

    void build(int idx) {
        Node& n = nd(idx);
        string g = gp(n);
        OpRec o = blank_op(0, idx);
        const string& t = n.type;
        if (t == "Distance3D") {
            need_args(n, 2);
            auto id = h5.ri(g + "/id");
            int m = (int)id.size() / 2;
            check_range(id, parent(n, 0).n_elem, n.name);
            alloc_coord(n, m, 1);
            o.kind = OP_DIST3D;
            o.ia[0] = T.ci_push(id);
            o.scratch = T.falloc(m * 3);
            o.ia[1] = T.falloc(m * 6);
            if (n.parents[0] == n.parents[1]) {
                o.plans[0] = atom_plan(parent(n, 0), id, 2, o.ia[1], 3);
            } else {
                vector<int> a(id.size(), -1), b(id.size(), -1);
                for (int e = 0; e < m; ++e) { a[2 * e] = id[2 * e]; b[2 * e + 1] = id[2 * e + 1]; }
                o.plans[0] = atom_plan(parent(n, 0), a, 2, o.ia[1], 3);
                o.plans[1] = atom_plan(parent(n, 1), b, 2, o.ia[1], 3);
            }
        } else if (t == "Angle") {
            need_args(n, 1);
            auto id = h5.ri(g + "/id");
            int m = (int)id.size() / 3;
            check_range(id, parent(n, 0).n_elem, n.name);
            alloc_coord(n, m, 1);
            o.kind = OP_ANGLE;
            o.ia[0] = T.ci_push(id);
            o.scratch = T.falloc(m * 6);
            o.ia[1] = T.falloc(m * 9);
            o.plans[0] = atom_plan(parent(n, 0), id, 3, o.ia[1], 3);
        } else if (t == "Dihedral") {
            need_args(n, 1);
            auto id = h5.rd(g + "/id");
            vector<int> ii(id.size());
            for (size_t k = 0; k < id.size(); ++k) ii[k] = (int)id[k];
            int m = (int)ii.size() / 4;
            vector<int> pl(ii.size(), -1);
            for (int e = 0; e < m; ++e) {
                bool dummy = ii[4 * e] < 0 || ii[4 * e + 3] < 0;
                for (int k = 0; k < 4; ++k) {
                    if (!dummy) {
                        if (ii[4 * e + k] >= parent(n, 0).n_elem) fail(n.name + " atom index out of range");
                        pl[4 * e + k] = ii[4 * e + k];
                    }
                }
            }
            alloc_coord(n, m, 1);
            o.kind = OP_DIHEDRAL;
            o.ia[0] = T.ci_push(ii);
            o.scratch = T.falloc(m * 12);
            o.ia[1] = T.falloc(m * 12);
            o.plans[0] = atom_plan(parent(n, 0), pl, 4, o.ia[1], 3);
        } else if (t == "Spring") {
            need_args(n, 1);
            n.potential = true;
            auto id = h5.ri(g + "/id");
            const Node& p = parent(n, 0);
            check_range(id, p.n_elem, n.name);
            o.kind = OP_SPRING;
            o.ia[0] = T.ci_push(id);
            o.ia[1] = T.cf_push(h5.rf(g + "/equil_dist"));
            o.ia[2] = T.cf_push(h5.rf(g + "/spring_const"));
            o.ia[3] = h5.attr_i(g, "dim1", 0, true);
            o.ia[4] = h5.attr_i(g, "pbc", 0, true);
            o.ia[5] = (int)id.size();
            o.fa[0] = h5.attr_f(g, "box_len");
            if (o.ia[3] >= p.width) fail(n.name + ": dim1 beyond argument width");
            o.scratch = T.falloc(id.size());
            vector<std::pair<int, int>> con;
            for (size_t e = 0; e < id.size(); ++e) con.push_back({id[e], (int)e});
            o.plans[0] = T.plan(p.n_elem, 1, p.sens, p.width, o.ia[3], o.scratch, 1, con);
        } else if (t == "rama_coord") {
            need_args(n, 2);
            auto i1 = h5.ri(g + "/index_pos1");
            auto i2 = h5.ri(g + "/index_pos2");
            check_range(i1, parent(n, 0).n_elem, n.name);
            check_range(i2, parent(n, 1).n_elem, n.name);
            alloc_coord(n, (int)i1.size(), 2);
            o.kind = OP_RAMA_COORD;
            o.ia[0] = T.ci_push(i1);
            o.ia[1] = T.ci_push(i2);
            vector<std::pair<int, int>> c0, c1;
            for (size_t e = 0; e < i1.size(); ++e) { c0.push_back({i1[e], (int)(2 * e)}); c1.push_back({i2[e], (int)(2 * e + 1)}); }
            const Node &p0 = parent(n, 0), &p1 = parent(n, 1);
            o.plans[0] = T.plan(p0.n_elem, 1, p0.sens, p0.width, 0, n.sens, 1, c0);
            o.plans[1] = T.plan(p1.n_elem, 1, p1.sens, p1.width, 0, n.sens, 1, c1);
        } else if (t == "cat_pos") {
            need_args(n, 2);
            auto i1 = h5.ri(g + "/index_pos1");
            auto i2 = h5.ri(g + "/index_pos2");
            int w = h5.attr_i(g, "n_dim", 0, true);
            const Node &p0 = parent(n, 0), &p1 = parent(n, 1);
            check_range(i1, p0.n_elem, n.name);
            check_range(i2, p1.n_elem, n.name);
            if (w > p0.width || w > p1.width) fail(n.name + ": n_dim beyond argument width");
            alloc_coord(n, (int)(i1.size() + i2.size()), w);
            o.kind = OP_CAT_POS;
            vector<int> all(i1);
            all.insert(all.end(), i2.begin(), i2.end());
            o.ia[0] = T.ci_push(all);
            o.ia[1] = (int)i1.size();
            vector<std::pair<int, int>> c0, c1;
            for (size_t e = 0; e < i1.size(); ++e) c0.push_back({i1[e], (int)e});
            for (size_t e = 0; e < i2.size(); ++e) c1.push_back({i2[e], (int)(i1.size() + e)});
            o.plans[0] = T.plan(p0.n_elem, w, p0.sens, p0.width, 0, n.sens, w, c0);
            o.plans[1] = T.plan(p1.n_elem, w, p1.sens, p1.width, 0, n.sens, w, c1);
        } else if (t == "rama_map_pot") {
            need_args(n, 1);
            n.potential = true;
            auto res = h5.ri(g + "/residue_id");
            auto mid = h5.ri(g + "/rama_map_id");
            auto sh = h5.shape(g + "/rama_pot");
            if (sh.size() != 3 || sh[1] != sh[2]) fail(n.name + ": rama_pot must be (L,n,n)");
            const Node& p = parent(n, 0);
            check_range(res, p.n_elem, n.name);
            check_range(mid, (int)sh[0], n.name);
            auto raw = h5.rd(g + "/rama_pot");
            o.kind = OP_RAMA_MAP_POT;
            o.ia[0] = T.ci_push(res);
            o.ia[1] = T.ci_push(mid);
            o.ia[2] = T.cf_push(periodic_bicubic((int)sh[0], (int)sh[1], (int)sh[2], raw));
            o.ia[3] = (int)sh[1];
            o.ia[4] = (int)res.size();
            o.fa[0] = (float)sh[1] * (0.5f / 3.141592653589793f - 1e-7f);
            o.fa[1] = 3.141592653589793f;
            o.ia[5] = h5.attr_i(g, "log_pot", 1, false);  // per-residue potential logged (src/rama_map_pot.cpp:34,50)
            o.scratch = T.falloc(res.size() * 3);
            vector<std::pair<int, int>> con;
            for (size_t e = 0; e < res.size(); ++e) con.push_back({res[e], (int)e});
            o.plans[0] = T.plan(p.n_elem, 2, p.sens, p.width, 0, o.scratch, 2, con);
        } else if (t == "affine_alignment") {
            need_args(n, 1);
            auto atoms = h5.ri(g + "/atoms");
            auto ref = h5.rf(g + "/ref_geom");
            int m = (int)atoms.size() / 3;
            check_range(atoms, parent(n, 0).n_elem, n.name);
            alloc_coord(n, m, 7);
            o.kind = OP_AFFINE;
            o.ia[0] = T.ci_push(atoms);
            o.ia[1] = T.cf_push(ref);
            o.scratch = T.falloc(m * 20);
            o.ia[2] = T.falloc(m * 9);
            o.plans[0] = atom_plan(parent(n, 0), atoms, 3, o.ia[2], 3);
        } else if (t == "placement_scalar" || t == "placement_fixed_point_vector_only" ||
                   t == "placement_fixed_point_vector_scalar") {
            auto ares_d = h5.rd(g + "/affine_residue");
            vector<int> ares(ares_d.size());
            for (size_t k = 0; k < ares.size(); ++k) ares[k] = (int)ares_d[k];
            auto layer = h5.ri(g + "/layer_index");
            auto sh = h5.shape(g + "/placement_data");
            const Node& aff = parent(n, 0);
            check_range(ares, aff.n_elem, n.name);
            o.kind = OP_PLACEMENT;
            o.ia[0] = T.ci_push(ares);
            o.ia[1] = T.ci_push(layer);
            int m = (int)layer.size();
            if (t == "placement_scalar") {
                need_args(n, 2);
                if (sh.size() != 4 || sh[3] != 1) fail(n.name + ": placement_data must be (L,nx,ny,1)");
                check_range(layer, (int)sh[0], n.name);
                auto rres = h5.ri(g + "/rama_residue");
                check_range(rres, parent(n, 1).n_elem, n.name);
                auto raw = h5.rd(g + "/placement_data");
                alloc_coord(n, m, 1);
                o.ia[2] = T.cf_push(periodic_bicubic((int)sh[0], (int)sh[1], (int)sh[2], raw));
                o.ia[3] = 1;
                o.ia[4] = T.ci_push(rres);
                o.ia[5] = (int)sh[1];
                o.ia[6] = (int)sh[2];
                o.ia[7] = 1;
                o.fa[0] = (float)sh[1] * (0.5f / 3.141592653589793f - 1e-7f);
                o.fa[1] = (float)sh[2] * (0.5f / 3.141592653589793f - 1e-7f);
                o.fa[2] = 3.141592653589793f;
                o.scratch = T.falloc(m * 2);
                o.ia[8] = T.falloc(m * 2);
                const Node& rp = parent(n, 1);
                vector<std::pair<int, int>> con;
                for (int e = 0; e < m; ++e) con.push_back({rres[e], e});
                o.plans[1] = T.plan(rp.n_elem, 2, rp.sens, rp.width, 0, o.ia[8], 2, con);
            } else {
                need_args(n, 1);
                int dim = t == "placement_fixed_point_vector_only" ? 6 : 7;
                if (sh.size() != 2 || (int)sh[1] != dim) fail(n.name + ": placement_data width mismatch");
                check_range(layer, (int)sh[0], n.name);
                alloc_coord(n, m, dim);
                o.ia[2] = T.cf_push(h5.rf(g + "/placement_data"));
                o.ia[3] = 0;
                o.ia[7] = dim;
                o.ia[8] = T.falloc(m * 6);
                vector<std::pair<int, int>> con;
                for (int e = 0; e < m; ++e) con.push_back({ares[e], e});
                o.plans[0] = T.plan(aff.n_elem, 6, aff.sens, aff.width, 0, o.ia[8], 6, con);
            }
        } else if (t == "infer_H_O") {
            need_args(n, 1);
            auto did = h5.ri(g + "/donors/id");
            auto aid = h5.ri(g + "/acceptors/id");
            auto dbl = h5.rf(g + "/donors/bond_length");
            auto abl = h5.rf(g + "/acceptors/bond_length");
            vector<int> id(did);
            id.insert(id.end(), aid.begin(), aid.end());
            vector<float> bl(dbl);
            bl.insert(bl.end(), abl.begin(), abl.end());
            check_range(id, parent(n, 0).n_elem, n.name);
            int m = (int)bl.size();
            alloc_coord(n, m, 6);
            o.kind = OP_INFER_HO;
            o.ia[0] = T.ci_push(id);
            o.ia[1] = T.cf_push(bl);
            o.scratch = T.falloc(m * 12);
            o.ia[2] = T.falloc(m * 9);
            o.plans[0] = atom_plan(parent(n, 0), id, 3, o.ia[2], 3);
        } else if (t == "protein_hbond") {
            need_args(n, 1);
            o.kind = OP_PROTEIN_HBOND;
            o.graph = graph(n, g, EDGE_PROTEIN_HBOND, false, &parent(n, 0), &parent(n, 0));
            const PairGraph& G = T.graphs[o.graph];
            alloc_coord(n, G.n1 + G.n2, 7);
            o.scratch = T.falloc(G.n1 + G.n2);
        } else if (t == "hbond_coverage") {
            need_args(n, 2);
            o.kind = OP_HBOND_COVERAGE;
            o.graph = graph(n, g, EDGE_HBOND_COVERAGE, false, &parent(n, 0), &parent(n, 1));
            alloc_coord(n, T.graphs[o.graph].n2, 1);
        } else if (t == "hbond_energy") {
            need_args(n, 2);
            n.potential = true;
            build_hbond_energy(n, g, o);
        } else if (t == "rotamer") {
            if (n.parents.size() < 2 || n.parents.size() > 5) fail(n.name + ": needs bead node and 1-4 energy nodes");
            n.potential = true;
            build_rotamer(n, g, o);
        } else if (t == "weighted_pos") {
            need_args(n, 2);
            auto ip = h5.ri(g + "/index_pos");
            auto iw = h5.ri(g + "/index_weight");
            const Node &p0 = parent(n, 0), &p1 = parent(n, 1);
            check_range(ip, p0.n_elem, n.name);
            check_range(iw, p1.n_elem, n.name);
            int m = (int)ip.size();
            alloc_coord(n, m, 4);
            o.kind = OP_WEIGHTED_POS;
            o.ia[0] = T.ci_push(ip);
            o.ia[1] = T.ci_push(iw);
            o.scratch = T.falloc(m);
            vector<std::pair<int, int>> c0, c1;
            for (int e = 0; e < m; ++e) { c0.push_back({ip[e], e}); c1.push_back({iw[e], e}); }
            o.plans[0] = T.plan(p0.n_elem, 3, p0.sens, p0.width, 0, n.sens, 4, c0);
            o.plans[1] = T.plan(p1.n_elem, 1, p1.sens, p1.width, 0, o.scratch, 1, c1);
        } else if (t == "environment_coverage") {
            need_args(n, 2);
            o.kind = OP_ENV_COVERAGE;
            o.graph = graph(n, g, EDGE_ENV_WEIGHTED, false, &parent(n, 0), &parent(n, 1));
            int naa = h5.attr_i(g, "num_aa_types", 0, true);
            auto aa = h5.ri(g + "/aa_types");
            const PairGraph& G = T.graphs[o.graph];
            const int* id2 = T.ci.data() + G.id2;
            int mx = (int)aa.size();
            for (int j = 0; j < G.n2; ++j) {
                if (id2[j] < 0) fail(n.name + ": negative id2");
                mx = std::max(mx, id2[j] + 1);
            }
            for (int x : aa)
                if (x < 0 || x >= naa) fail(n.name + ": aa type out of range");
            aa.resize(mx, 0);
            if (naa < 1 || naa * NT > LDS_FLOATS) fail(n.name + ": num_aa_types out of range");
            o.ia[0] = naa;
            o.ia[1] = T.ci_push(aa);
            alloc_coord(n, G.n1 * naa, 1);
        } else if (t == "hb_environment_coverage") {
            need_args(n, 2);
            o.kind = OP_HB_ENV_COVERAGE;
            o.graph = graph(n, g, EDGE_ENV_HBOND, false, &parent(n, 0), &parent(n, 1));
            alloc_coord(n, T.graphs[o.graph].n1, 1);
        } else if (t == "sigmoid_coupling") {
            need_args(n, 1);
            n.potential = true;
            const Node& p = parent(n, 0);
            if (p.width != 1) fail(n.name + ": argument width must be 1");
            auto scale = h5.rf(g + "/scale");
            int nt = (int)scale.size();
            auto ct = h5.ri(g + "/coupling_types");
            auto W = h5.rf(g + "/weights");
            if ((int)W.size() < nt * nt) fail(n.name + ": weights too short");
            check_range(ct, nt, n.name);
            if ((int)ct.size() * nt > p.n_elem) fail(n.name + ": coupling types exceed input");
            o.kind = OP_SIGMOID_COUPLING;
            o.ia[0] = p.n_elem / nt;
            o.ia[1] = T.ci_push(ct);
            o.ia[2] = T.cf_push(W);
            o.ia[3] = T.cf_push(scale);
            o.ia[4] = T.cf_push(h5.rf(g + "/center"));
            o.ia[5] = T.cf_push(h5.rf(g + "/sharpness"));
            o.ia[6] = nt;
            o.ia[7] = h5.attr_i(g, "number_independent_weights", 1, false);  // for get_param_deriv only
        } else if (t == "bb_sigmoid_coupling") {
            need_args(n, 2);
            n.potential = true;
            const Node &p0 = parent(n, 0), &p1 = parent(n, 1);
            int nt = h5.attr_i(g, "num_aa_types", 0, true);
            auto W = h5.rf(g + "/weights");
            if ((int)W.size() < nt) fail(n.name + ": weights too short");
            if (p0.width != 1 || p1.width != 1) fail(n.name + ": argument widths must be 1");
            o.kind = OP_BB_SIGMOID_COUPLING;
            o.ia[0] = nt;
            o.ia[1] = T.cf_push(W);
            o.ia[2] = p0.n_elem / nt;
            if (o.ia[2] > p1.n_elem) fail(n.name + ": backbone coverage shorter than environment");
            o.fa[0] = h5.attr_f(g, "scale");
            o.fa[1] = h5.attr_f(g, "center");
            o.fa[2] = h5.attr_f(g, "sharpness");
            o.fa[3] = h5.attr_f(g, "hbond_weight");
        } else if (t == "backbone_pairs") {
            need_args(n, 1);
            n.potential = true;
            build_backbone(n, g, o);
        } else {
            fail("No node type found for name '" + n.name + "'");
        }
        if (n.potential) n.pot_slot = T.n_pot++;
        o.n_elem = n.n_elem;
        o.width = n.width;
        o.out = n.out;
        o.sens = n.sens;
        o.pot_slot = n.pot_slot;
        set_inputs(o, n);
        n.op = (int)T.ops.size();
        T.ops.push_back(o);
    }

    // Periodic bicubic coefficients (L,nx,ny,4,4): tensor product of periodic
    // interpolating cubic splines through data at integer grid points, solved
    // in double (the unique spline of solve_periodic_2d_spline, src/spline.cpp:262-292).
    static vector<double> periodic_1d(const vector<double>& d) {
        int n = (int)d.size();
        // cyclic tridiagonal system (s[i-1]+4 s[i]+s[i+1])/6 = d[i] by Sherman-Morrison
        vector<double> a(n, 1. / 6.), b(n, 2. / 3.), c(n, 1. / 6.);
        auto thomas = [&](vector<double> rhs, vector<double> bb) {
            vector<double> cc(c);
            for (int k = 1; k < n; ++k) {
                double m = a[k] / bb[k - 1];
                bb[k] -= m * cc[k - 1];
                rhs[k] -= m * rhs[k - 1];
            }
            rhs[n - 1] /= bb[n - 1];
            for (int k = n - 2; k >= 0; --k) rhs[k] = (rhs[k] - cc[k] * rhs[k + 1]) / bb[k];
            return rhs;
        };
        double gamma = -b[0];
        vector<double> bb(b);
        bb[0] -= gamma;
        bb[n - 1] -= a[0] * c[n - 1] / gamma;
        vector<double> u(n, 0.);
        u[0] = gamma;
        u[n - 1] = c[n - 1];
        vector<double> y = thomas(d, bb), z = thomas(u, bb);
        double fact = (y[0] + a[0] * y[n - 1] / gamma) / (1. + z[0] + a[0] * z[n - 1] / gamma);
        vector<double> s(n);
        for (int i = 0; i < n; ++i) s[i] = y[i] - fact * z[i];
        vector<double> coef(4 * n);
        for (int i = 0; i < n; ++i) {
            double sm = s[(i - 1 + n) % n], s0 = s[i], s1 = s[(i + 1) % n], s2 = s[(i + 2) % n];
            coef[4 * i + 0] = (sm + 4 * s0 + s1) / 6.;
            coef[4 * i + 1] = (s1 - sm) / 2.;
            coef[4 * i + 2] = (sm - 2 * s0 + s1) / 2.;
            coef[4 * i + 3] = (-sm + 3 * s0 - 3 * s1 + s2) / 6.;
        }
        return coef;
    }
    static vector<float> periodic_bicubic(int L, int nx, int ny, const vector<double>& raw) {
        vector<float> out((size_t)L * nx * ny * 16);
        for (int l = 0; l < L; ++l) {
            vector<double> ay((size_t)nx * ny * 4);
            for (int ix = 0; ix < nx; ++ix) {
                vector<double> row(raw.begin() + ((size_t)l * nx + ix) * ny, raw.begin() + ((size_t)l * nx + ix + 1) * ny);
                auto c = periodic_1d(row);
                std::copy(c.begin(), c.end(), ay.begin() + (size_t)ix * ny * 4);
            }
            for (int iy = 0; iy < ny; ++iy)
                for (int py = 0; py < 4; ++py) {
                    vector<double> col(nx);
                    for (int ix = 0; ix < nx; ++ix) col[ix] = ay[((size_t)ix * ny + iy) * 4 + py];
                    auto c = periodic_1d(col);
                    for (int ix = 0; ix < nx; ++ix)
                        for (int px = 0; px < 4; ++px)
                            out[((((size_t)l * nx + ix) * ny + iy) * 16) + px * 4 + py] = (float)c[4 * ix + px];
                }
        }
        return out;
    }

    int cap(int n, int want) { return std::max(4, std::min(n, (int)(want * T.cap_scale))); }

    // InteractionGraph (src/interaction_graph.h:262-405) lowering.
    int graph(const Node& n, const string& g, int kind, bool sym, const Node* p1, const Node* p2) {
        PairGraph G;
        std::memset(&G, 0, sizeof(G));
        G.kind = kind;
        G.sym = sym;
        if (sym) p2 = p1;
        string s1 = sym ? "" : "1";
        auto loc1 = h5.ri(g + "/index" + s1);
        auto ty1 = h5.ri(g + "/type" + s1);
        auto idd1 = h5.rd(g + "/id" + s1);
        vector<int> id1(idd1.size());
        for (size_t k = 0; k < id1.size(); ++k) id1[k] = (int)idd1[k];
        vector<int> loc2, ty2, id2;
        if (sym) {
            loc2 = loc1; ty2 = ty1; id2 = id1;
        } else {
            loc2 = h5.ri(g + "/index2");
            ty2 = h5.ri(g + "/type2");
            auto idd2 = h5.rd(g + "/id2");
            for (double x : idd2) id2.push_back((int)x);
        }
        if (ty1.size() != loc1.size() || id1.size() != loc1.size() || ty2.size() != loc2.size() ||
            id2.size() != loc2.size())
            fail(n.name + ": index/type/id size mismatch");
        check_range(loc1, p1->n_elem, n.name);
        check_range(loc2, p2->n_elem, n.name);
        auto sh = h5.shape(g + "/interaction_param");
        if (sh.size() != 3) fail(n.name + ": interaction_param must be 3-d");
        auto param = h5.rf(g + "/interaction_param");
        int nt1 = (int)sh[0], nt2 = (int)sh[1], np = (int)sh[2];
        check_range(ty1, nt1, n.name);
        check_range(ty2, nt2, n.name);
        // within one side each producer row may be fed by one element only
        for (auto* L : {&loc1, &loc2}) {
            vector<int> s(*L);
            std::sort(s.begin(), s.end());
            if (std::adjacent_find(s.begin(), s.end()) != s.end()) fail(n.name + ": repeated producer index");
        }
        G.n1 = (int)loc1.size();
        G.n2 = (int)loc2.size();
        {   // the two gathers may share a phase when they never touch the same producer row
            bool same_node = sym || p1 == p2;
            std::set<int> r1(loc1.begin(), loc1.end());
            bool shared = false;
            for (int j : loc2) shared = shared || r1.count(j);
            G.gather_split = (same_node && shared) ? 1 : 0;
        }
        G.loc1 = T.ci_push(loc1); G.loc2 = T.ci_push(loc2);
        G.type1 = T.ci_push(ty1); G.type2 = T.ci_push(ty2);
        G.id1 = T.ci_push(id1); G.id2 = T.ci_push(id2);
        G.param = T.cf_push(param, 64);
        G.n_type1 = nt1;
        G.n_type2 = nt2;
        G.n_param = np;
        G.w1 = p1->width; G.w2 = p2->width;
        G.out1 = p1->out; G.out2 = p2->out;
        G.sens1 = p1->sens; G.sens2 = p2->sens;
        float cutoff = 0.f, buffer = 0.f;
        int want = 64;  // initial slot capacity per row (grown on overflow)
        switch (kind) {
        case EDGE_ROTAMER: {
            G.dim1 = G.dim2 = 6; G.excl = -1;
            float inv_dx = 1.f / 0.7f;
            cutoff = (float)((12 - 2 - 1e-6) / inv_dx);
            if (np != 54) fail(n.name + ": rotamer pair table must have 54 parameters");
            want = 64;
        } break;
        case EDGE_HBOND_COVERAGE: {
            G.dim1 = 7; G.dim2 = 6; G.excl = 2;
            float inv_dx = 1.f / 0.625f;
            cutoff = (float)((10 - 2 - 1e-6) / inv_dx);
            if (np != 50) fail(n.name + ": coverage table must have 50 parameters");
            want = 64;
        } break;
        case EDGE_PROTEIN_HBOND:
            G.dim1 = G.dim2 = 6; G.excl = 1;
            cutoff = sqrtf(3.5f * 3.5f);
            if (np != 8) fail(n.name + ": hbond table must have 8 parameters");
            want = 16;
            break;
        case EDGE_ENV_WEIGHTED:
        case EDGE_ENV_HBOND:
            G.dim1 = 6; G.dim2 = kind == EDGE_ENV_WEIGHTED ? 4 : 6; G.excl = 2;
            if (np != 4) fail(n.name + ": coverage table must have 4 parameters");
            for (int k = 0; k < nt1 * nt2; ++k) cutoff = std::max(cutoff, param[k * 4] + 1.f / param[k * 4 + 1]);
            want = kind == EDGE_ENV_WEIGHTED ? 256 : 32;
            break;
        }
        if (p1->width < G.dim1 || p2->width < G.dim2) fail(n.name + ": argument narrower than interaction");
        buffer = 1.0f + 0.2f * cutoff;
        G.ddim1 = G.dim1; G.ddim2 = G.dim2;
        long table = (long)nt1 * nt2 * np + 64;
        G.param_len = table <= LDS_FLOATS ? (int)table : 0;
        finish_graph(G, cutoff, buffer, want);
        T.graphs.push_back(G);
        T.graph_node.push_back((int)(&n - T.nodes.data()));
        return (int)T.graphs.size() - 1;
    }

    // The Verlet buffer only trades cache rebuild frequency against cache
    // size: the edge set is the exact refine of the cache whatever its radius.
    // The reference uses 1 + 0.2*cutoff (src/interaction_graph.h:397); this
    // engine scales it by UPSIDE_BUFFER_SCALE (default 1).
    void finish_graph(PairGraph& G, float cutoff, float buffer, int want) {
        if (const char* e = std::getenv("UPSIDE_BUFFER_SCALE")) {
            double f = std::atof(e);
            if (f >= 0.25 && f <= 8.0) {
                buffer = (float)(buffer * f);
                want = (int)(want * std::pow((cutoff + buffer) / (cutoff + buffer / f), 3.0) + 1);
            }
        }
        G.cutoff2 = cutoff * cutoff;
        float cc = cutoff + buffer;
        G.cache_cut2 = cc * cc;
        G.cache_cut = cc;
        G.drift2 = (0.5f * buffer) * (0.5f * buffer);
        long full = G.sym ? (long)G.n1 * (G.n1 - 1) / 2 : (long)G.n1 * G.n2;
        long ecap = std::min(full, (long)(std::max(G.n1, 1) * (double)want * T.cap_scale));
        G.ecap = (int)std::max(ecap, 16L);
        int npack = G.sym ? G.n1 : G.n1 + G.n2;
        G.cache_pos = T.falloc(npack * 3);
        G.pos_pack = T.falloc(npack * 3);
        long slots = G.ecap;
        G.off = T.ialloc(G.n1 + 1);
        G.nbr = T.ialloc(slots);
        G.row = T.ialloc(slots);
        G.toff = T.ialloc(G.n2 + 1);
        G.tslot = T.ialloc(slots);
        G.tmp = T.ialloc(slots);
        G.tcur = T.ialloc(std::max(G.n1, G.n2) + 1);
        G.val = T.falloc(slots);
        G.d1 = T.falloc(slots * G.ddim1);
        G.d2 = T.falloc(slots * G.ddim2);
        G.esens = T.falloc(slots);
        G.act = T.ialloc(slots);
        G.state = T.ialloc(4);
    }

    void build_backbone(Node& n, const string& g, OpRec& o) {
        const Node& aff = parent(n, 0);
        auto res = h5.ri(g + "/id");
        auto nat = h5.ri(g + "/n_atom");
        auto ref = h5.rf(g + "/ref_pos");
        int m = (int)res.size();
        check_range(res, aff.n_elem, n.name);
        if ((int)nat.size() != m || (int)ref.size() != m * 12) fail(n.name + ": size mismatch");
        float mx = 0.f;
        for (int r = 0; r < m; ++r) {
            if (nat[r] < 0 || nat[r] > 4) fail(n.name + ": n_atom out of range");
            for (int a = 0; a < nat[r]; ++a) {
                const float* q = &ref[(r * 4 + a) * 3];
                float m2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
                mx = std::max(m2 * (1.f / sqrtf(m2)), mx);
            }
        }
        for (auto& x : ref)
            if (std::isnan(x)) x = 0.f;
        float cutoff = 2 * mx + sqrtf(3.f * 3.f + 0.1f * 3.f);
        PairGraph G;
        std::memset(&G, 0, sizeof(G));
        G.kind = EDGE_BACKBONE;
        G.sym = 1;
        G.n1 = G.n2 = m;
        G.dim1 = G.dim2 = 3;
        G.ddim1 = G.ddim2 = 6;
        G.w1 = G.w2 = aff.width;
        G.out1 = G.out2 = aff.out;
        G.sens1 = G.sens2 = aff.sens;
        G.loc1 = G.loc2 = T.ci_push(res);
        vector<int> zeros(m, 0);
        G.type1 = G.type2 = T.ci_push(zeros);
        G.id1 = G.id2 = T.ci_push(res);
        G.param = T.cf_push(vector<float>(1, 0.f), 64);
        G.n_type1 = 1;
        G.n_type2 = 1;
        G.n_param = 0;
        G.param_len = 0;
        G.excl = 1;
        finish_graph(G, cutoff, 1.0f, 32);
        o.kind = OP_BACKBONE_PAIRS;
        o.ia[0] = G.loc1;
        o.ia[1] = T.cf_push(ref);
        o.ia[2] = T.ci_push(nat);
        T.graphs.push_back(G);
        T.graph_node.push_back((int)(&n - T.nodes.data()));
        o.graph = (int)T.graphs.size() - 1;
    }

    void build_hbond_energy(Node& n, const string& g, OpRec& o) {
        const Node& pp = parent(n, 0);
        const Node& rama = parent(n, 1);
        auto params = h5.rf(g + "/parameters");
        auto dres = h5.ri(g + "/donor_resid");
        auto ares = h5.ri(g + "/acceptor_resid");
        int n_res = (int)h5.shape(g + "/rama_resid")[0];
        int nd = (int)dres.size(), na = (int)ares.size();
        if (params.size() < 12) fail(n.name + ": 12 parameters expected");
        if (nd + na != pp.n_elem || pp.width < 7) fail(n.name + ": donors+acceptors must match protein_hbond");
        if (rama.n_elem < n_res || rama.width < 2) fail(n.name + ": rama node too small");
        check_range(dres, n_res, n.name);
        check_range(ares, n_res, n.name);
        vector<vector<int>> h1(n_res), h2(std::max(n_res - 1, 0));
        vector<int> fd(n_res, -1), fa(n_res, -1);
        for (int i = 0; i < nd; ++i) {
            h1[dres[i]].push_back(i);
            if (dres[i] - 1 >= 0 && dres[i] - 1 < n_res - 1) h2[dres[i] - 1].push_back(i);
            if (fd[dres[i]] < 0) fd[dres[i]] = i;
        }
        for (int i = 0; i < na; ++i) {
            h1[ares[i]].push_back(nd + i);
            if (ares[i] < n_res - 1) h2[ares[i]].push_back(nd + i);
            if (fa[ares[i]] < 0) fa[ares[i]] = nd + i;
        }
        auto csr = [&](const vector<vector<int>>& L, vector<int>& off, vector<int>& idx) {
            off.assign(1, 0);
            for (auto& l : L) { idx.insert(idx.end(), l.begin(), l.end()); off.push_back((int)idx.size()); }
        };
        vector<int> o1, i1, o2, i2;
        csr(h1, o1, i1);
        csr(h2, o2, i2);
        o.kind = OP_HBOND_ENERGY;
        o.ia[0] = n_res;
        o.ia[1] = T.ci_push(o1);
        o.ia[2] = T.ci_push(i1);
        o.ia[3] = T.ci_push(o2);
        o.ia[4] = T.ci_push(i2);
        o.ia[5] = T.ci_push(fd);
        o.ia[6] = T.ci_push(fa);
        for (int k = 0; k < 12; ++k) o.fa[k] = params[k];
    }

    void build_rotamer(Node& n, const string& g, OpRec& o) {
        const Node& beads = parent(n, 0);
        o.kind = OP_ROTAMER;
        o.graph = graph(n, g + "/pair_interaction", EDGE_ROTAMER, true, &beads, nullptr);
        PairGraph& G = T.graphs[o.graph];
        const int* ids = T.ci.data() + G.id1;
        int nb = G.n1;
        vector<int> node_of(nb), nrot, bead0, brot(nb);
        std::map<int, int> key_node;
        for (int b = 0; b < nb; ++b) {
            int id = ids[b];
            int rot = id & 15, nr = (id >> 4) & 15;
            if (rot >= nr || !(nr == 1 || nr == 3 || nr == 6))
                fail(n.name + ": invalid rotamer id (rotamer counts 1, 3, 6 supported)");
            int key = id >> 4;
            auto it = key_node.find(key);
            int node;
            if (it == key_node.end()) {
                node = (int)nrot.size();
                key_node[key] = node;
                nrot.push_back(nr);
                bead0.push_back(b);
            } else {
                node = it->second;
                if (node != (int)nrot.size() - 1) fail(n.name + ": beads of a residue must be contiguous");
            }
            node_of[b] = node;
            brot[b] = rot;
        }
        bead0.push_back(nb);
        for (size_t a = 0; a < nrot.size(); ++a) {
            if (bead0[a + 1] - bead0[a] != nrot[a])
                fail(n.name + ": a residue must have one bead per rotamer state");
            for (int k = 0; k < nrot[a]; ++k)
                if (brot[bead0[a] + k] != k) fail(n.name + ": rotamer beads must be listed in state order");
        }
        RotamerDev R;
        std::memset(&R, 0, sizeof(R));
        R.n_node = (int)nrot.size();
        R.node_nrot = T.ci_push(nrot);
        R.node_bead0 = T.ci_push(bead0);
        R.bead_rot = T.ci_push(brot);
        R.bead_node = T.ci_push(node_of);
        R.n_prob = (int)n.parents.size() - 1;
        for (int q = 0; q < R.n_prob; ++q) {
            const Node& p = parent(n, q + 1);
            if (p.n_elem != beads.n_elem)
                fail(n.name + ": energy node " + p.name + " must have one element per bead position");
            R.prob_out[q] = p.out;
            R.prob_sens[q] = p.sens;
            R.prob_w[q] = p.width;
        }
        R.damping = h5.attr_f(g, "damping");
        R.max_iter = h5.attr_i(g, "max_iter", 0, true);
        R.tol = h5.attr_f(g, "tol");
        R.chunk = h5.attr_i(g, "iteration_chunk_size", 0, true);
        if (R.chunk < 1) fail(n.name + ": iteration_chunk_size must be positive");
        int nn = R.n_node;
        if (nn >= 65536) fail(n.name + ": too many residue nodes (BP edge endpoints are packed in 16 bits)");
        R.ccap = (int)std::max(16.0, nn * 64.0 * T.cap_scale);
        long cc = R.ccap;
        R.coff = T.ialloc(nn + 1);
        R.ca = T.ialloc(cc);
        R.cb = T.ialloc(cc);
        R.slot_cand = T.ialloc(G.ecap);
        R.ioff = T.ialloc(nn + 1);
        R.iinc = T.ialloc(2 * cc);
        R.eidx = T.ialloc(cc + 1);
        R.bpc = T.ialloc(cc);
        R.bpf = T.ialloc(cc);
        R.bps = T.ialloc(cc);
        R.ninc = T.ialloc(nn);
        R.binc = T.ialloc(2 * cc);
        R.P = T.falloc(cc * 36);
        R.msg = T.falloc(cc * 12);
        R.msg2 = T.falloc(cc * 12);
        R.marg = T.falloc(cc * 36);
        R.e1x = T.falloc(nb);
        R.ncen = T.falloc(nn * 4);
        R.nprob = T.falloc(nn * 6);
        R.ncur = T.falloc(nn * 6);
        R.nold = T.falloc(nn * 6);
        R.noff = T.falloc(nn);
        R.feat = T.falloc(nn * MLP_IN);
        R.nprev = T.falloc(nn * 6);
        R.stats = T.ialloc(BP_HIST + BP_HIST_BINS);
        o.ia[0] = T.ci_push(node_of);
        T.rots.push_back(R);
        o.rot = (int)T.rots.size() - 1;
    }
};

static void load_topology(Topology& T, H5& h5, int n_atom) {
    T.n_atom = n_atom;
    T.nf = T.ni = 0;
    T.cf.clear();
    T.ci.clear();
    T.nodes.clear();
    T.ops.clear();
    T.graphs.clear();
    T.rots.clear();
    T.graph_node.clear();
    T.sens_zero.clear();
    T.n_pot = 0;
    string pg = "/input/potential";
    std::map<string, vector<string>> deps;
    deps["pos"] = {};
    for (auto& nm : h5.children(pg)) deps[nm] = h5.attr_strings(pg + "/" + nm, "arguments");
    for (auto& kv : deps)
        for (auto& a : kv.second)
            if (!deps.count(a))
                fail("Node " + kv.first + " takes " + a + " as an argument, but no node of that name can be found.");
    vector<string> order;
    std::set<string> done;
    std::map<string, bool> active;
    for (auto& kv : deps) active[kv.first] = true;
    for (size_t round = 0; round < deps.size(); ++round)
        for (auto& kv : deps) {
            if (!active[kv.first]) continue;
            bool ok = true;
            for (auto& a : kv.second)
                if (!done.count(a)) { ok = false; break; }
            if (ok) { order.push_back(kv.first); done.insert(kv.first); active[kv.first] = false; }
        }
    for (auto& kv : active)
        if (kv.second) fail("Unsatisfiable dependency " + kv.first + " in potential computation");
    // pos node first (DerivEngine ctor), then the others in topological order
    Node pos;
    pos.name = "pos";
    pos.type = "pos";
    T.nodes.push_back(pos);
    Builder B(T, h5);
    B.alloc_coord(T.nodes[0], n_atom, 3);
    std::map<string, int> index;
    index["pos"] = 0;
    for (auto& nm : order) {
        if (nm == "pos") continue;
        Node n;
        n.name = nm;
        n.type = node_type(nm);
        if (n.type.empty()) fail("No node type found for name '" + nm + "'");
        n.args = deps[nm];
        for (auto& a : n.args) {
            int p = index.at(a);
            if (T.nodes[p].out < 0)
                fail(a + " is not an intermediate value, but it is an argument of " + nm);
            n.parents.push_back(p);
        }
        index[nm] = (int)T.nodes.size();
        T.nodes.push_back(n);
        for (int p : T.nodes.back().parents) T.nodes[p].children.push_back(index[nm]);
        try {
            B.build((int)T.nodes.size() - 1);
        } catch (std::exception& e) {
            fail("while adding '" + nm + "', " + e.what());
        }
    }
    // execution levels (src/deriv_engine.cpp:127-178)
    auto& N = T.nodes;
    N[0].germ_level = 0;
    T.germ.assign(1, 0);
    T.deriv.clear();
    for (int lvl = 1;; ++lvl) {
        bool unfinished = false;
        for (int i = 0; i < (int)N.size(); ++i) {
            auto& n = N[i];
            if (n.germ_level == -1) {
                unfinished = true;
                bool ok = true;
                for (int p : n.parents) if (N[p].germ_level == -1 || N[p].germ_level == lvl) ok = false;
                if (ok) { n.germ_level = lvl; T.germ.push_back(i); }
            }
            if (n.deriv_level == -1 && n.germ_level != -1 && i != 0) {
                unfinished = true;
                bool ok = true;
                for (int ch : n.children) if (N[ch].deriv_level == -1 || N[ch].deriv_level == lvl) ok = false;
                if (ok) { n.deriv_level = lvl; T.deriv.push_back(i); }
            }
        }
        if (!unfinished) break;
        if (lvl > 10000) fail("cyclic node graph");
    }
    T.pot = T.falloc(std::max(T.n_pot, 1));
    T.mom = T.falloc(n_atom * 3);
    T.err = T.ialloc(2 + MAX_LANES);  // the team's word, then one per section sub-team
    T.bar = T.ialloc(BAR_WORDS);
    // two buffers each: the team, sub-teams 1..G-1, 0..G/2-1, G/2..G-1, the four quarters
    T.tred = T.falloc(16 * MAX_TEAM);
    vector<int> z;
    for (auto& r : T.sens_zero) { z.push_back(r.first); z.push_back(r.second); }
    T.szero = T.ialloc(z.size());
    T.sens_zero_host = z;
}

}  // namespace upside

// ============================================================== runtime ===
namespace upside {

// Host Threefry-4x32-20 (same algorithm as the device one) for replica
// exchange decisions (RandomGenerator, src/random.h:19-67).
static void threefry_host(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4]) {
    static const int R0[8] = {10, 11, 13, 23, 6, 17, 25, 18};
    static const int R1[8] = {26, 21, 27, 5, 20, 11, 10, 20};
    auto rotl = [](uint32_t v, int r) { return (v << r) | (v >> (32 - r)); };
    uint32_t ks[5] = {key[0], key[1], key[2], key[3], 0x1BD11BDAu ^ key[0] ^ key[1] ^ key[2] ^ key[3]};
    uint32_t x[4] = {ctr[0] + ks[0], ctr[1] + ks[1], ctr[2] + ks[2], ctr[3] + ks[3]};
    for (int r = 0; r < 20; ++r) {
        if ((r & 1) == 0) {
            x[0] += x[1]; x[1] = rotl(x[1], R0[r & 7]) ^ x[0];
            x[2] += x[3]; x[3] = rotl(x[3], R1[r & 7]) ^ x[2];
        } else {
            x[0] += x[3]; x[3] = rotl(x[3], R0[r & 7]) ^ x[0];
            x[2] += x[1]; x[1] = rotl(x[1], R1[r & 7]) ^ x[2];
        }
        if ((r & 3) == 3) {
            int s = (r >> 2) + 1;
            for (int i = 0; i < 4; ++i) x[i] += ks[(s + i) % 5];
            x[3] += (uint32_t)s;
        }
    }
    for (int i = 0; i < 4; ++i) out[i] = x[i];
}

static inline float u01_open_closed(uint32_t x) {  // r123::u01<float>: (0, 1]
    return (float)x * 2.3283064365386962890625e-10f + 1.16415321826934814453125e-10f;
}

// Monte Carlo pivot moves (PivotSampler, src/monte_carlo_sampler.cpp:19-137,
// input group /input/pivot_moves written by py/upside_config.py:2543-2552):
// a residue's (phi, psi) is redrawn from its Rama proposal map and everything
// downstream of it is rotated rigidly.
struct PivotMoves {
    int n_layer = 0, n_bin = 0;
    vector<int> atom;     // [n_loc][5] prevC N CA C nextN
    vector<int> range;    // [n_loc][2] atoms rotated
    vector<int> restype;  // [n_loc] proposal layer
    vector<float> lprob;  // proposal_pot + log(sum exp(-proposal_pot)) per layer
    vector<float> cdf;    // cumulative proposal probability per layer
    int n_loc() const { return (int)restype.size(); }

    void load(const H5& h5, int n_atom) {
        const string g = "/input/pivot_moves";
        if (!h5.exists(g)) return;
        auto sh = h5.shape(g + "/proposal_pot");
        if (sh.size() != 3 || sh[1] != sh[2]) fail("pivot_moves: proposal_pot must be (n_layer, n_bin, n_bin)");
        n_layer = (int)sh[0];
        n_bin = (int)sh[1];
        atom = h5.ri(g + "/pivot_atom");
        range = h5.ri(g + "/pivot_range");
        restype = h5.ri(g + "/pivot_restype");
        const int nl = n_loc();
        if ((int)atom.size() != 5 * nl || (int)range.size() != 2 * nl) fail("pivot_moves: inconsistent sizes");
        for (int k = 0; k < nl; ++k) {
            if (restype[k] < 0 || restype[k] >= n_layer) fail("invalid pivot restype");
            for (int a = 0; a < 5; ++a) {
                if (atom[5 * k + a] < 0 || atom[5 * k + a] >= n_atom) fail("pivot_atom out of range");
                if (range[2 * k] <= atom[5 * k + a] && atom[5 * k + a] < range[2 * k + 1])
                    fail("pivot_range cannot contain any atoms in pivot_atom");
            }
            if (range[2 * k] < 0 || range[2 * k + 1] > n_atom) fail("pivot_range out of range");
        }
        auto pot = h5.rf(g + "/proposal_pot");
        const int nn = n_bin * n_bin;
        lprob = pot;
        cdf.assign(pot.size(), 0.f);
        for (int l = 0; l < n_layer; ++l) {  // normalisation in double (:70-83)
            double sum = 0.0;
            for (int i = 0; i < nn; ++i) {
                sum += std::exp(-(double)pot[l * nn + i]);
                cdf[l * nn + i] = (float)sum;
            }
            const double inv = 1.0 / sum, lsum = std::log(sum);
            for (int i = 0; i < nn; ++i) {
                cdf[l * nn + i] = (float)((double)cdf[l * nn + i] * inv);
                lprob[l * nn + i] = (float)((double)lprob[l * nn + i] + lsum);
            }
            cdf[(l + 1) * nn - 1] = 1.f;
        }
    }

    static void rot(float* U, float angle, const float* a) {  // axis_angle_to_rot, src/affine.h:49-64
        const float c = std::cos(angle), s = std::sin(angle), C = 1.f - c;
        const float x = a[0], y = a[1], z = a[2];
        U[0] = x * x * C + c;     U[1] = x * y * C - z * s; U[2] = x * z * C + y * s;
        U[3] = y * x * C + z * s; U[4] = y * y * C + c;     U[5] = y * z * C - x * s;
        U[6] = z * x * C - y * s; U[7] = z * y * C + x * s; U[8] = z * z * C + c;
    }
    static float dihedral(const float* r1, const float* r2, const float* r3, const float* r4) {
        // Blondel-Karplus form of dihedral_germ (src/vector_math.h:766-796)
        float F[3], G[3], H[3], A[3], B[3], C[3];
        for (int k = 0; k < 3; ++k) { F[k] = r1[k] - r2[k]; G[k] = r2[k] - r3[k]; H[k] = r4[k] - r3[k]; }
        auto cross = [](const float* u, const float* v, float* w) {
            w[0] = u[1] * v[2] - u[2] * v[1]; w[1] = u[2] * v[0] - u[0] * v[2]; w[2] = u[0] * v[1] - u[1] * v[0];
        };
        cross(F, G, A);
        cross(H, G, B);
        cross(B, A, C);
        const float gmag = std::sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
        return std::atan2(C[0] * G[0] + C[1] * G[1] + C[2] * G[2], (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]) * gmag);
    }

    // Propose a move on pos (n_atom x 3) from four uniforms u (the first draw
    // of the move's generator); returns log(proposal ratio) (:89-137).
    float propose(float* pos, const float* u) const {
        const float PI = 3.14159265358979323846f;
        int loc = (int)(n_loc() * u[2]);
        if (loc == n_loc()) loc--;
        const int* at = &atom[5 * loc];
        const int l = restype[loc], nn = n_bin * n_bin;
        const int bin = (int)(std::lower_bound(cdf.begin() + l * nn, cdf.begin() + (l + 1) * nn, u[3]) -
                              (cdf.begin() + l * nn));
        const float new_lp = lprob[l * nn + bin];
        const int phi_bin = bin / n_bin, psi_bin = bin % n_bin;
        const float w = 2.f * PI / n_bin;
        const float new_phi = w * (phi_bin + u[0] - 0.5f) - PI;
        const float new_psi = w * (psi_bin + u[1] - 0.5f) - PI;
        const float* prevC = pos + 3 * at[0];
        const float* N = pos + 3 * at[1];
        const float* CA = pos + 3 * at[2];
        const float* Cp = pos + 3 * at[3];
        const float* nextN = pos + 3 * at[4];
        const float old_phi = dihedral(prevC, N, CA, Cp), old_psi = dihedral(N, CA, Cp, nextN);
        int ob_phi = (int)((old_phi + PI) * (0.5f / PI) * n_bin + 0.5f);
        int ob_psi = (int)((old_psi + PI) * (0.5f / PI) * n_bin + 0.5f);
        ob_phi = ob_phi >= n_bin ? 0 : ob_phi;
        ob_psi = ob_psi >= n_bin ? 0 : ob_psi;
        const float old_lp = lprob[(l * n_bin + ob_phi) * n_bin + ob_psi];
        float phi_o[3], psi_o[3], ax1[3], ax2[3];
        for (int k = 0; k < 3; ++k) { phi_o[k] = CA[k]; psi_o[k] = Cp[k]; ax1[k] = CA[k] - N[k]; ax2[k] = Cp[k] - CA[k]; }
        const float n1 = 1.f / std::sqrt(ax1[0] * ax1[0] + ax1[1] * ax1[1] + ax1[2] * ax1[2]);
        const float n2 = 1.f / std::sqrt(ax2[0] * ax2[0] + ax2[1] * ax2[1] + ax2[2] * ax2[2]);
        for (int k = 0; k < 3; ++k) { ax1[k] *= n1; ax2[k] *= n2; }
        float Uphi[9], Upsi[9];
        rot(Uphi, new_phi - old_phi, ax1);
        rot(Upsi, new_psi - old_psi, ax2);
        auto move = [&](int a) {
            float* y = pos + 3 * a;
            float d[3], t[3];
            for (int k = 0; k < 3; ++k) d[k] = y[k] - psi_o[k];
            for (int k = 0; k < 3; ++k) t[k] = psi_o[k] + (Upsi[3 * k] * d[0] + Upsi[3 * k + 1] * d[1] + Upsi[3 * k + 2] * d[2]);
            for (int k = 0; k < 3; ++k) d[k] = t[k] - phi_o[k];
            for (int k = 0; k < 3; ++k) y[k] = phi_o[k] + (Uphi[3 * k] * d[0] + Uphi[3 * k + 1] * d[1] + Uphi[3 * k + 2] * d[2]);
        };
        move(at[3]);
        move(at[4]);
        for (int a = range[2 * loc]; a < range[2 * loc + 1]; ++a) move(a);
        return new_lp - old_lp;
    }
};

// Op kinds with work in the sensitivity pass (the others push in their value pass).
static bool has_deriv_step(int kind) {
    switch (kind) {
    case OP_DIST3D: case OP_ANGLE: case OP_DIHEDRAL: case OP_RAMA_COORD: case OP_CAT_POS: case OP_AFFINE:
    case OP_PLACEMENT: case OP_INFER_HO: case OP_PROTEIN_HBOND: case OP_HBOND_COVERAGE: case OP_WEIGHTED_POS:
    case OP_ENV_COVERAGE: case OP_HB_ENV_COVERAGE:
        return true;
    default:
        return false;
    }
}

struct Program {
    string path;
    Topology T;
    OpRec* d_ops = nullptr;
    PairGraph* d_graphs = nullptr;
    RotamerDev* d_rots = nullptr;
    float* d_cf = nullptr;
    int* d_ci = nullptr;
    std::map<int, vector<float>> param_set;  // parameters replaced by set_param, by node
    PivotMoves pivots;                       // /input/pivot_moves, if present
    vector<OpRec> prog;                      // the uploaded program (d_ops)
    vector<int> prog_node;                   // node of each program entry

    // Parameter vector of a node in the reference's get_param layout, and
    // where it lives here: a pair graph's interaction table (constant floats);
    // sigmoid coupling scale, center, sharpness per type then the weights
    // (src/environment.cpp:925-939); backbone sigmoid coupling scale, center,
    // sharpness, hbond weight (:1102-1109) and hbond energy's 12 parameters
    // (src/hbond.cpp:653) in the op record.  Empty: none.
    vector<float> get_params(int node) const {
        const Node& n = T.nodes[node];
        for (size_t g = 0; g < T.graphs.size(); ++g)
            if (T.graph_node[g] == node) {
                const PairGraph& G = T.graphs[g];
                return vector<float>(T.cf.begin() + G.param,
                                     T.cf.begin() + G.param + (size_t)G.n_type1 * G.n_type2 * G.n_param);
            }
        if (n.op < 0) return {};
        const OpRec& o = T.ops[n.op];
        vector<float> v;
        if (o.kind == OP_SIGMOID_COUPLING) {
            const int nt = o.ia[6];
            for (int q : {3, 4, 5}) v.insert(v.end(), T.cf.begin() + o.ia[q], T.cf.begin() + o.ia[q] + nt);
            v.insert(v.end(), T.cf.begin() + o.ia[2], T.cf.begin() + o.ia[2] + nt * nt);
        } else if (o.kind == OP_BB_SIGMOID_COUPLING) {
            v.assign(o.fa, o.fa + 4);
        } else if (o.kind == OP_HBOND_ENERGY) {
            v.assign(o.fa, o.fa + 12);
        }
        return v;
    }
    // Write a parameter vector of that layout back (sizes checked by the caller).
    void put_params(int node, const vector<float>& v) {
        const Node& n = T.nodes[node];
        for (size_t g = 0; g < T.graphs.size(); ++g)
            if (T.graph_node[g] == node) {
                std::copy(v.begin(), v.end(), T.cf.begin() + T.graphs[g].param);
                return;
            }
        OpRec& o = T.ops[n.op];
        if (o.kind == OP_SIGMOID_COUPLING) {
            const int nt = o.ia[6];
            size_t k = 0;
            for (int q : {3, 4, 5})
                for (int i = 0; i < nt; ++i) T.cf[o.ia[q] + i] = v[k++];
            for (int i = 0; i < nt * nt; ++i) T.cf[o.ia[2] + i] = v[k++];
        } else {
            for (size_t i = 0; i < v.size(); ++i) o.fa[i] = v[i];
        }
    }

    void release() {
        for (void* p : {(void*)d_ops, (void*)d_graphs, (void*)d_rots, (void*)d_cf, (void*)d_ci})
            if (p) (void)hipFree(p);
        d_ops = nullptr; d_graphs = nullptr; d_rots = nullptr; d_cf = nullptr; d_ci = nullptr;
    }
    ~Program() { release(); }

    // Value-pass order for two-lane sections (no counterpart in the reference,
    // whose value pass runs its exec levels in order, src/deriv_engine.cpp:127-178).
    // Any topological order computes the same values; only the order in which
    // potential nodes add into their parents' sensitivities changes, which
    // moves those sums by rounding.
    //  1. Elementwise ops whose ancestors are all elementwise come first (a
    //     stable partition, which keeps both parts in topological order): one
    //     run of them, then the pair-graph ops.
    //  2. From the first pair-graph op on, ops are dealt to two lanes: an op
    //     that descends from an op of one lane, adds into the same
    //     sensitivities as a potential of that lane, or shares its pair graph,
    //     joins that lane; a free op joins the lane with less work (pair-graph
    //     ops weighted by kind, lane_weight; elementwise ops 1).  The rotamer
    //     op, ops tied to both lanes, elementwise ops after the rotamer and
    //     every descendant of these are deferred past the section, in their
    //     order.  The two halves of a team run the two lanes at once.
    vector<int> lane_of;                        // node -> lane (0 outside sections)
    int n_lanes = MAX_LANES;                    // lanes per section (UPSIDE_LANES=2: two)
    int rot_graph_node = -1;                    // rotamer whose pair graph lane 0 of the section refreshes
    vector<std::pair<int, int>> lane_sections;  // [first op, end op) in program order
    // rough relative cost of a pair-graph op by its edge kind (a 500-residue
    // chain: environment graphs ~140 us per step, coverage ~90, backbone and
    // hbond ~50, hbond environment ~35)
    int lane_weight(int g) const {
        switch (T.graphs[g].kind) {
        case EDGE_ROTAMER: return 14;
        case EDGE_ENV_WEIGHTED: return 14;
        case EDGE_HBOND_COVERAGE: return 9;
        case EDGE_ENV_HBOND: return 4;
        default: return 5;
        }
    }
    void order_germ_for_lanes() {
        const int nn = (int)T.nodes.size();
        lane_of.assign(nn, 0);
        lane_sections.clear();
        auto elementwise = [&](int i) {
            const OpRec& o = T.ops[T.nodes[i].op];
            return o.graph < 0 && o.rot < 0;
        };
        // anc[i][q]: node i descends from node q (pos, node 0, left out)
        vector<vector<char>> anc(nn, vector<char>(nn, 0));
        for (size_t k = 1; k < T.germ.size(); ++k) {
            const int i = T.germ[k];
            for (int p : T.nodes[i].parents) {
                if (p == 0) continue;
                anc[i][p] = 1;
                for (int q = 0; q < nn; ++q) anc[i][q] |= anc[p][q];
            }
        }
        vector<int> first, rest;
        for (size_t k = 1; k < T.germ.size(); ++k) {
            const int i = T.germ[k];
            bool pure = elementwise(i);
            for (int q = 0; q < nn && pure; ++q)
                if (anc[i][q] && !elementwise(q)) pure = false;
            (pure ? first : rest).push_back(i);
        }
        const int NL = n_lanes;
        vector<vector<int>> in_lane(NL);
        vector<int> deferred, work(NL, 0);
        // the first rotamer's pair-list refresh and edge evaluation read only
        // its bead placements: when those run before the section, lane 0
        // starts with them
        rot_graph_node = -1;
        for (int i : rest) {
            const OpRec& o = T.ops[T.nodes[i].op];
            if (o.kind != OP_ROTAMER) continue;
            const PairGraph& G = T.graphs[o.graph];
            bool ready = true;
            for (int p : T.nodes[i].parents)
                if (T.nodes[p].out == G.out1 || T.nodes[p].out == G.out2)
                    ready = ready && std::find(first.begin(), first.end(), p) != first.end();
            if (ready) {
                rot_graph_node = i;
                work[0] += lane_weight(o.graph);
            }
            break;
        }
        auto tied = [&](int i, const vector<int>& lane) {
            const Node& ni = T.nodes[i];
            const OpRec& oi = T.ops[ni.op];
            for (int j : lane) {
                if (anc[i][j]) return true;
                if (oi.graph >= 0 && oi.graph == T.ops[T.nodes[j].op].graph) return true;
                if (ni.potential && T.nodes[j].potential)
                    for (int p : ni.parents)
                        for (int q : T.nodes[j].parents)
                            if (p == q) return true;
            }
            return false;
        };
        // lane of op i: the one lane it is tied to, the least loaded lane
        // (lowest index on ties) when it is tied to none, -1 (deferred past
        // the section) when tied to two or more
        auto pick = [&](int i, const vector<vector<int>>& lanes, const vector<int>& w) {
            int tl = -1, nt = 0;
            for (int l = 0; l < NL; ++l)
                if (tied(i, lanes[l])) { tl = l; ++nt; }
            if (nt > 1) return -1;
            if (nt == 1) return tl;
            int best = 0;
            for (int l = 1; l < NL; ++l)
                if (w[l] < w[best]) best = l;
            return best;
        };
        // the elementwise ops themselves form a first section: one workgroup
        // of each sub-team runs its lane.  The affine alignment, whose
        // eigen-solve costs about eight of the others, starts lane 1; the rest
        // are dealt in order as above, one unit each.
        vector<vector<int>> fl(NL);
        vector<int> fdef;
        {
            vector<int> fw(NL, 0);
            for (int i : first)
                if (T.ops[T.nodes[i].op].kind == OP_AFFINE && fl[1].empty()) {
                    fl[1].push_back(i);
                    fw[1] = 8;
                    lane_of[i] = 1;
                }
            for (int i : first) {
                if (!fl[1].empty() && i == fl[1][0]) continue;
                int lane = pick(i, fl, fw);
                for (int j : fdef) lane = anc[i][j] ? -1 : lane;
                if (lane < 0) { fdef.push_back(i); continue; }
                fl[lane].push_back(i);
                fw[lane] += 1;
                lane_of[i] = lane;
            }
            int busy = 0;
            for (int l = 0; l < NL; ++l) busy += fl[l].size() >= 2;
            if (busy < 2) {
                for (int l = 0; l < NL; ++l) for (int i : fl[l]) lane_of[i] = 0;
                for (auto& v : fl) v.clear();
                fdef.clear();
            }
        }
        // pair-graph ops after the first rotamer op that do not depend on it
        // join the section too; elementwise ops after it stay behind, beside
        // the solve (measured on a 500-residue chain: 2 % faster than moving
        // all, 2 % faster than moving none)
        bool after_rot = false;
        size_t r = 0;
        for (; r < rest.size(); ++r) {
            const int i = rest[r];
            const bool rot = T.ops[T.nodes[i].op].kind == OP_ROTAMER;
            if (rot && after_rot) break;  // one rotamer per section
            // tied to two lanes, a rotamer, or descending from a deferred op:
            // after the section
            const int lane = rot ? -1 : pick(i, in_lane, work);
            bool defer = rot || lane < 0 || (after_rot && elementwise(i));
            for (int j : deferred) defer = defer || anc[i][j];
            if (rot) after_rot = true;
            if (defer) { deferred.push_back(i); continue; }
            in_lane[lane].push_back(i);
            work[lane] += elementwise(i) ? 1 : lane_weight(T.ops[T.nodes[i].op].graph);
            lane_of[i] = lane;
        }
        // a section needs pair-graph work in two lanes at least
        auto has_pair = [&](const vector<int>& lane) {
            for (int i : lane)
                if (!elementwise(i)) return true;
            return false;
        };
        int pair_lanes = 0;
        for (int l = 0; l < NL; ++l) pair_lanes += has_pair(in_lane[l]);
        T.germ.assign(1, 0);
        bool any_fl = false;
        for (auto& v : fl) any_fl = any_fl || !v.empty();
        if (any_fl) {  // the elementwise section, then its deferred ops
            for (int i : first)
                if (std::find(fdef.begin(), fdef.end(), i) == fdef.end()) T.germ.push_back(i);
            lane_sections.push_back({0, (int)T.germ.size() - 1});
            T.germ.insert(T.germ.end(), fdef.begin(), fdef.end());
        } else {
            T.germ.insert(T.germ.end(), first.begin(), first.end());
        }
        if (pair_lanes < 2) {
            for (int l = 0; l < NL; ++l) for (int i : in_lane[l]) lane_of[i] = 0;
            rot_graph_node = -1;
            T.germ.insert(T.germ.end(), rest.begin(), rest.end());
            return;
        }
        // the section's ops in their program order, then the deferred ops, then the rest
        const int k0 = (int)first.size();
        for (size_t q = 0; q < r; ++q)
            if (std::find(deferred.begin(), deferred.end(), rest[q]) == deferred.end()) T.germ.push_back(rest[q]);
        lane_sections.push_back({k0, (int)T.germ.size() - 1});
        T.germ.insert(T.germ.end(), deferred.begin(), deferred.end());
        T.germ.insert(T.germ.end(), rest.begin() + r, rest.end());
    }

    void load(int n_atom, double cap_scale, bool upload = true) {
        H5 h5(path);
        if (n_atom <= 0) n_atom = (int)h5.shape("/input/pos")[0];
        T.cap_scale = cap_scale;
        load_topology(T, h5, n_atom);
        pivots = PivotMoves();
        pivots.load(h5, n_atom);
        for (auto& kv : param_set) put_params(kv.first, kv.second);  // set_param survives a reload
        release();
        // potential nodes push their sensitivities in the value pass: their
        // deriv steps have nothing to do and stay out of the program
        T.deriv.erase(std::remove_if(T.deriv.begin(), T.deriv.end(),
                                     [&](int i) { return !has_deriv_step(T.ops[T.nodes[i].op].kind); }),
                      T.deriv.end());
        const char* so = std::getenv("UPSIDE_SOLO_OPS");
        const bool solo_ok = !(so && so[0] == '0');
        const char* ln = std::getenv("UPSIDE_LANES");
        n_lanes = (ln && ln[0] == '2') ? 2 : MAX_LANES;
        if (solo_ok && !(ln && ln[0] == '0')) order_germ_for_lanes();
        else lane_of.assign(T.nodes.size(), 0), lane_sections.clear(), rot_graph_node = -1;
        prog.clear();
        prog_node.clear();
        for (size_t k = 1; k < T.germ.size(); ++k) {
            prog.push_back(T.ops[T.nodes[T.germ[k]].op]);
            prog_node.push_back(T.germ[k]);
        }
        for (int i : T.deriv) {
            prog.push_back(T.ops[T.nodes[i].op]);
            prog_node.push_back(i);
        }
        // ops without a pair graph or rotamer solve are elementwise over at most
        // a few thousand elements: one workgroup runs them (UPSIDE_SOLO_OPS=0:
        // the whole team)
        for (auto& o : prog) o.solo = (solo_ok && o.graph < 0 && o.rot < 0) ? 1 : 0;
        for (size_t k = 0; k + 1 < T.germ.size(); ++k) {
            prog[k].lane = lane_of[T.germ[k + 1]];
            prog[k].lane_end = 0;
        }
        for (auto& o : prog) o.rot_graph_op = o.rot_graph_lane = -1;
        for (auto& sec : lane_sections) prog[sec.first].lane_end = sec.second;
        if (rot_graph_node >= 0) {  // lane 0 of the pair-graph section (the last) refreshes the rotamer graph
            const auto& sec = lane_sections.back();
            for (int q = sec.second; q < (int)T.germ.size() - 1; ++q)
                if (prog_node[q] == rot_graph_node) {
                    prog[sec.first].rot_graph_op = q;
                    prog[sec.first].rot_graph_lane = 0;
                    break;
                }
        }
        // A rotamer node has no coordinate output, so the germ ops after it
        // usually do not depend on it: those that neither descend from it nor
        // push sensitivities into its inputs run beside its BP solve
        // (UPSIDE_OVERLAP=0: never).
        const char* ov = std::getenv("UPSIDE_OVERLAP");
        const bool overlap_ok = !(ov && ov[0] == '0');
        const int ng = (int)T.germ.size() - 1;
        for (int k = 0; k < ng && overlap_ok; ++k) {
            if (prog[k].kind != OP_ROTAMER) continue;
            const int rn = T.germ[k + 1];
            std::set<int> inputs(T.nodes[rn].parents.begin(), T.nodes[rn].parents.end());
            auto descends = [&](int x) {  // x has the rotamer node among its ancestors
                vector<int> stack(T.nodes[x].parents.begin(), T.nodes[x].parents.end());
                std::set<int> seen;
                while (!stack.empty()) {
                    int y = stack.back();
                    stack.pop_back();
                    if (y == rn) return true;
                    if (!seen.insert(y).second) continue;
                    for (int z : T.nodes[y].parents) stack.push_back(z);
                }
                return false;
            };
            int j = k + 1;
            for (; j < ng; ++j) {
                const Node& nj = T.nodes[T.germ[j + 1]];
                if (prog[j].kind == OP_ROTAMER || descends(T.germ[j + 1])) break;
                bool pushes_into_inputs = false;
                if (nj.potential)
                    for (int p : nj.parents) pushes_into_inputs = pushes_into_inputs || inputs.count(p);
                if (pushes_into_inputs) break;
            }
            prog[k].overlap_end = j;  // may be k + 1: an empty range, kept if deriv ops run beside the solve
        }
        // Deriv ops whose sensitivities are complete once the ops before the
        // rotamer solve and its overlap ops have run (every consumer is a
        // potential among those, other than the rotamer, or itself such a
        // deriv op) run beside the solve, after the overlap ops, in deriv
        // order (UPSIDE_EARLY_DERIV=0: never).  The solve writes no
        // sensitivities, and the rotamer's own push into its inputs comes
        // after the join, so nothing they touch is touched at the same time.
        for (auto& o : prog) o.early = 0;
        const char* ed = std::getenv("UPSIDE_EARLY_DERIV");
        const bool early_ok = overlap_ok && !(ed && ed[0] == '0');
        for (int k = 0; k < ng && early_ok; ++k) {
            if (prog[k].kind != OP_ROTAMER || prog[k].overlap_end <= k) continue;
            const int e = prog[k].overlap_end;
            vector<int> germ_pos(T.nodes.size(), -1);
            for (int q = 0; q < ng; ++q) germ_pos[T.germ[q + 1]] = q;
            vector<char> early(T.nodes.size(), 0);
            for (size_t d = 0; d < T.deriv.size(); ++d) {
                const int i = T.deriv[d];
                bool ok = true;
                for (int ch : T.nodes[i].children) {
                    if (T.nodes[ch].potential)
                        ok = ok && germ_pos[ch] >= 0 && germ_pos[ch] < e && germ_pos[ch] != k;
                    else
                        ok = ok && early[ch];
                }
                if (ok) {
                    early[i] = 1;
                    prog[ng + d].early = 1;
                }
            }
            break;  // beside the first rotamer solve only
        }
        // a rotamer op with neither overlap ops nor early deriv ops runs whole
        for (int k = 0; k < ng; ++k) {
            if (prog[k].kind != OP_ROTAMER || prog[k].overlap_end != k + 1) continue;
            bool any = false;
            for (size_t d = 0; d < T.deriv.size(); ++d) any = any || prog[ng + d].early;
            if (!any) prog[k].overlap_end = 0;
        }
        if (!upload) return;  // lowering only (check_program)
        auto up = [](void** dst, const void* src, size_t bytes, const char* what) {
            hip_check(hipMalloc(dst, std::max(bytes, (size_t)16)), what);
            if (bytes) hip_check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
        };
        up((void**)&d_ops, prog.data(), prog.size() * sizeof(OpRec), "program");
        up((void**)&d_graphs, T.graphs.data(), T.graphs.size() * sizeof(PairGraph), "graphs");
        up((void**)&d_rots, T.rots.data(), T.rots.size() * sizeof(RotamerDev), "rotamers");
        up((void**)&d_cf, T.cf.data(), T.cf.size() * sizeof(float), "constants");
        up((void**)&d_ci, T.ci.data(), T.ci.size() * sizeof(int), "indices");
    }
    int n_germ() const { return (int)T.germ.size() - 1; }
    int n_deriv() const { return (int)T.deriv.size(); }

    // Invariants of the lowered program that the kernel relies on; "" when
    // they all hold, else the first violation.
    //  1. the value pass (germ) is a topological order of every node with an
    //     op, each once, after node 0;
    //  2. in a section, ops of different lanes are independent: neither
    //     descends from the other, they share no pair graph, and no two
    //     potentials push into a shared input;
    //  3. the section's rotamer pair-list refresh reads only nodes placed
    //     before the section;
    //  4. the ops of a rotamer's overlap range neither descend from the
    //     rotamer nor (potentials) push into its inputs;
    //  5. the deriv pass runs a node after every non-potential child's deriv
    //     op, and an early deriv op's consumers are potentials run before
    //     the solve (or early deriv ops themselves).
    string check_program() const {
        const int nn = (int)T.nodes.size(), ng = n_germ();
        if (T.germ.empty() || T.germ[0] != 0) return "value pass does not start at node 0";
        vector<int> pos(nn, -1);
        for (int q = 0; q <= ng; ++q) {
            const int i = T.germ[q];
            if (i < 0 || i >= nn || pos[i] >= 0) return "value pass repeats or leaves the nodes";
            pos[i] = q;
        }
        for (int i = 1; i < nn; ++i) {
            if (T.nodes[i].op >= 0 && pos[i] < 0) return "node " + T.nodes[i].name + " missing from the value pass";
            for (int p : T.nodes[i].parents)
                if (pos[i] >= 0 && (pos[p] < 0 || pos[p] >= pos[i]))
                    return "node " + T.nodes[i].name + " runs before its input " + T.nodes[p].name;
        }
        // transitive ancestry
        vector<vector<char>> anc(nn, vector<char>(nn, 0));
        for (int q = 1; q <= ng; ++q) {
            const int i = T.germ[q];
            for (int p : T.nodes[i].parents) {
                anc[i][p] = 1;
                for (int z = 0; z < nn; ++z) anc[i][z] |= anc[p][z];
            }
        }
        auto share_input = [&](int a, int b) {
            for (int p : T.nodes[a].parents)
                for (int q : T.nodes[b].parents)
                    if (p == q && p != 0) return true;
            return false;
        };
        for (auto& sec : lane_sections) {
            for (int x = sec.first; x < sec.second; ++x)
                for (int y = x + 1; y < sec.second; ++y) {
                    if (prog[x].lane == prog[y].lane) continue;
                    const int a = prog_node[x], b = prog_node[y];
                    const string pair = T.nodes[a].name + " / " + T.nodes[b].name;
                    if (anc[a][b] || anc[b][a]) return "lanes of a section depend on each other: " + pair;
                    if (prog[x].graph >= 0 && prog[x].graph == prog[y].graph) return "lanes share a pair graph: " + pair;
                    if (T.nodes[a].potential && T.nodes[b].potential && share_input(a, b))
                        return "lanes push into a shared input: " + pair;
                }
            for (int x = sec.first; x < sec.second; ++x)
                if (prog[x].lane < 0 || prog[x].lane >= n_lanes) return "lane out of range";
            const int rgo = prog[sec.first].rot_graph_op;
            if (rgo >= 0) {
                const PairGraph& G = T.graphs[prog[rgo].graph];
                for (int p : T.nodes[prog_node[rgo]].parents)
                    if ((T.nodes[p].out == G.out1 || T.nodes[p].out == G.out2) && pos[p] - 1 >= sec.first)
                        return "rotamer pair-list refresh reads " + T.nodes[p].name + " of its own section";
            }
        }
        for (int k = 0; k < ng; ++k) {
            if (prog[k].kind != OP_ROTAMER || prog[k].overlap_end <= k + 1) continue;
            const int rn = prog_node[k];
            for (int j = k + 1; j < prog[k].overlap_end; ++j) {
                const int x = prog_node[j];
                if (anc[x][rn]) return "overlap op " + T.nodes[x].name + " descends from the rotamer";
                if (T.nodes[x].potential)
                    for (int p : T.nodes[x].parents)
                        for (int q : T.nodes[rn].parents)
                            if (p == q) return "overlap potential " + T.nodes[x].name + " pushes into a rotamer input";
            }
        }
        vector<int> dpos(nn, -1);
        for (size_t d = 0; d < T.deriv.size(); ++d) dpos[T.deriv[d]] = (int)d;
        int rot_k = -1, rot_end = -1;
        for (int k = 0; k < ng && rot_k < 0; ++k)
            if (prog[k].kind == OP_ROTAMER && prog[k].overlap_end > k) { rot_k = k; rot_end = prog[k].overlap_end; }
        for (size_t d = 0; d < T.deriv.size(); ++d) {
            const int i = T.deriv[d];
            for (int ch : T.nodes[i].children) {
                if (T.nodes[ch].potential) {
                    if (prog[ng + d].early && !(pos[ch] - 1 < rot_end && pos[ch] - 1 != rot_k))
                        return "early deriv op " + T.nodes[i].name + " before its consumer " + T.nodes[ch].name;
                    continue;
                }
                if (dpos[ch] < 0 || dpos[ch] >= (int)d)
                    return "deriv op " + T.nodes[i].name + " runs before its consumer " + T.nodes[ch].name;
                if (prog[ng + d].early && !prog[ng + dpos[ch]].early)
                    return "early deriv op " + T.nodes[i].name + " waits on a late one";
            }
        }
        return "";
    }
};

struct System {
    std::shared_ptr<Program> prog;
    float* arena = nullptr;
    int* iarena = nullptr;
    float temperature = 1.f;   // current (annealed) temperature
    float temp0 = 1.f;         // initial temperature
    void release() { (void)hipFree(arena); (void)hipFree(iarena); arena = nullptr; iarena = nullptr; }
};

// Place per-system teams t on the N_XCD XCDs (a team never spans two):
// largest team first (ties: system order) onto the least loaded XCD, each XCD
// holding at most per_xcd workgroups.  False if they do not fit; else, when
// map is given, the block map (RunArgs::blockmap): workgroup b runs on XCD
// b % N_XCD, each XCD's teams in system order, -1 for padding.
static bool plan_team_map(const vector<int>& t, int per_xcd, vector<int>* map) {
    vector<int> order(t.size());
    for (size_t s = 0; s < t.size(); ++s) order[s] = (int)s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t[a] > t[b]; });
    vector<int> load(N_XCD, 0), where(t.size(), 0);
    for (int s : order) {
        int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        if (t[s] < 1 || load[x] + t[s] > per_xcd) return false;
        where[s] = x;
        load[x] += t[s];
    }
    if (map) {
        const int slots = *std::max_element(load.begin(), load.end());
        map->assign((size_t)N_XCD * slots, -1);
        vector<int> fill(N_XCD, 0);
        for (size_t s = 0; s < t.size(); ++s) {
            const int x = where[s];
            for (int r = 0; r < t[s]; ++r) (*map)[(size_t)(fill[x] + r) * N_XCD + x] = ((int)s << 14) | (t[s] << 7) | r;
            fill[x] += t[s];
        }
    }
    return true;
}

// Team sizes in proportion to the systems' sizes (atom counts): starting from
// one workgroup each, the system with the most atoms per workgroup (ties: the
// lower index) gains a workgroup while the total stays within budget, no team
// exceeds MAX_TEAM and the teams still fit (plan_team_map).  Empty if even
// one workgroup per system does not fit.
static vector<int> plan_balanced_teams(const vector<int>& atoms, int budget, int per_xcd) {
    const int n = (int)atoms.size();
    vector<int> t(n, 1);
    if (n > budget || !plan_team_map(t, per_xcd, nullptr)) return {};
    for (long long total = n; total < budget; ++total) {
        int best = -1;
        double bw = -1.0;
        for (int s = 0; s < n; ++s) {
            if (t[s] >= MAX_TEAM) continue;
            const double w = (double)atoms[s] / t[s];
            if (w > bw) { bw = w; best = s; }
        }
        if (best < 0) break;
        t[best] += 1;
        if (!plan_team_map(t, per_xcd, nullptr)) { t[best] -= 1; break; }
    }
    return t;
}

struct Batch {
    vector<std::shared_ptr<Program>> progs;
    vector<System> sys;
    vector<SysDesc> desc;        // host mirror
    SysDesc* d_desc = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    float dt = 0.009f;
    int thermostat_interval = 1;
    uint32_t base_seed = 0;
    bool have_mom = false;
    double cap_scale = 1.0;
    uint64_t round_num = 0;      // rounds integrated since creation (src/main.cpp:920 round_num)
    float thermostat_timescale = 5.f;  // --thermostat-timescale (src/main.cpp:477, :775-781)
    // simulated annealing (--anneal-factor/-start/-end/-duration, src/main.cpp:593-612)
    int anneal = 0;
    double anneal_factor = 1.0, anneal_start = 0.0, anneal_end = -1.0, anneal_duration = 1.0;
    double anneal_temp(double T0, double time) const {  // the reference's anneal_temp, in double
        const double T1 = T0 * anneal_factor;
        if (time <= anneal_start) return T0;
        if (time > anneal_end) return T1;
        const double f = std::max(0.0, (time - anneal_start) / anneal_duration);
        const double s = std::sqrt(T0) * (1.0 - f) + std::sqrt(T1) * f;
        return s * s;
    }
    // temperature in effect after the thermostat of the round before round_num
    // The temperature the reference's loop holds after round_num rounds: the
    // one set at its last thermostat round nr (nr % thermostat_interval == 0),
    // the schedule at time 3 dt (nr + 1) (src/main.cpp:974-977).
    void sync_temperature() {
        if (!anneal || round_num == 0 || thermostat_interval <= 0) return;
        const uint64_t last = round_num - 1;
        const uint64_t nr = last / (uint64_t)thermostat_interval * (uint64_t)thermostat_interval;
        const float time = 3.f * dt * (float)(nr + 1);
        for (auto& s : sys) s.temperature = (float)anneal_temp((double)s.temp0, (double)time);
    }
    bool profile = false;
    vector<unsigned long long*> prof;
    // wall-clock ticks per op, fixed slots, then team barriers per op (+1: outside the ops)
    int prof_len(int k) const { return 2 * (sys[k].prog->n_germ() + sys[k].prog->n_deriv()) + PROF_EXTRA + 1; }

    void enable_profile(bool on) {
        for (auto p : prof) (void)hipFree(p);
        prof.assign(sys.size(), nullptr);
        profile = on;
        if (on)
            for (size_t k = 0; k < sys.size(); ++k) {
                size_t bytes = prof_len((int)k) * sizeof(unsigned long long);
                hip_check(hipMalloc((void**)&prof[k], bytes), "profile");
                hip_check(hipMemset(prof[k], 0, bytes), "profile");
            }
        pull_desc();
        for (size_t k = 0; k < sys.size(); ++k) desc[k].prof = on ? prof[k] : nullptr;
        push_desc();
    }

    string profile_report(int k) {
        if (!profile) return "profiling disabled\n";
        int n = prof_len(k);
        vector<unsigned long long> t(n);
        hip_check(hipMemcpy(t.data(), prof[k], n * sizeof(unsigned long long), hipMemcpyDeviceToHost), "profile");
        int rate_khz = 100000;
        (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0);
        double tick = 1.0 / (rate_khz * 1e3);
        const Topology& T = sys[k].prog->T;
        string out;
        char line[256];
        if (team_of(k) > 1) out += string("team of ") + std::to_string(team_of(k)) + " workgroups, " +
                             (team_local ? "on one XCD (L2-local barriers)\n" : "spread over XCDs (device-scope barriers)\n");
        int ng = sys[k].prog->n_germ();
        const int nops = ng + (int)T.deriv.size(), nb0 = nops + PROF_EXTRA;
        unsigned long long nbar = 0;
        for (int i = nb0; i < n; ++i) nbar += t[i];
        for (int i = 0; i < nb0; ++i) {
            string label;
            if (i < ng) {
                label = "value  " + T.nodes[T.germ[i + 1]].name;
                const Program& P = *sys[k].prog;
                for (auto& sec : P.lane_sections)  // ops run by one half of the team
                    if (i >= sec.first && i < sec.second && team_of(k) >= 4)
                        label += " [lane " + std::to_string(P.lane_of[T.germ[i + 1]]) + "]";
            }
            else if (i < ng + (int)T.deriv.size()) label = "deriv  " + T.nodes[T.deriv[i - ng]].name;
            else {
                static const char* ex[] = {"rotamer BP solve (in value)", "rotamer graph rebuild (in value)",
                                           "thermostat", "integrator update",
                                           "pair-list rebuilds, other graphs (in value)",
                                           "pair edge evaluation, all graphs (in value)",
                                           "pair sensitivity gathers, all graphs",
                                           "rotamer fill (in value)", "rotamer candidates (in rebuild)",
                                           "rebuild: count scans", "rebuild: fill scans",
                                           "rebuild: transposed slot mapping",
                                           "rebuild: LDS grid builds", "rebuild: grid scans",
                                           "BP: message passes", "BP: node passes", "BP: staging and marginals",
                                           "beside the solve: overlap and early deriv ops"};
                label = ex[i - ng - (int)T.deriv.size()];
            }
            if (i < nops)
                snprintf(line, sizeof(line), "%-48s %12.6f s %8llu team barriers\n", label.c_str(), t[i] * tick,
                         t[nb0 + i]);
            else
                snprintf(line, sizeof(line), "%-48s %12.6f s\n", label.c_str(), t[i] * tick);
            out += line;
        }
        snprintf(line, sizeof(line), "%-48s %12llu\n%-48s %12llu\n", "team barriers outside the ops", t[n - 1],
                 "team barriers in all (rank 0)", nbar);
        out += line;
        for (size_t g = 0; g < T.graphs.size(); ++g) {
            int st[4];
            hip_check(hipMemcpy(st, sys[k].iarena + T.graphs[g].state, sizeof(st), hipMemcpyDeviceToHost), "state");
            snprintf(line, sizeof(line), "rebuilds %-39s %12d\n", T.nodes[T.graph_node[g]].name.c_str(), st[2]);
            out += line;
        }
        return out;
    }

    ~Batch() {
        for (auto& s : sys) s.release();
        (void)hipFree(d_blockmap);
        for (auto p : prof) (void)hipFree(p);
        (void)hipFree(d_desc);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void alloc_system(System& s) {
        const Topology& T = s.prog->T;
        hip_check(hipMalloc((void**)&s.arena, (size_t)std::max(T.nf, 4) * sizeof(float)), "arena");
        hip_check(hipMalloc((void**)&s.iarena, (size_t)std::max(T.ni, 4) * sizeof(int)), "int arena");
        hip_check(hipMemset(s.arena, 0, (size_t)std::max(T.nf, 4) * sizeof(float)), "arena clear");
        hip_check(hipMemset(s.iarena, 0, (size_t)std::max(T.ni, 4) * sizeof(int)), "int arena clear");
        if (!T.sens_zero_host.empty())
            hip_check(hipMemcpy(s.iarena + T.szero, T.sens_zero_host.data(), T.sens_zero_host.size() * sizeof(int),
                                hipMemcpyHostToDevice), "sens ranges");
    }

    SysDesc make_desc(int k) {
        System& s = sys[k];
        const Program& P = *s.prog;
        const Topology& T = P.T;
        SysDesc d;
        std::memset(&d, 0, sizeof(d));
        d.arena = s.arena;
        d.iarena = s.iarena;
        d.ops = P.d_ops;
        d.n_germ = P.n_germ();
        d.n_deriv = P.n_deriv();
        d.graphs = P.d_graphs;
        d.rots = P.d_rots;
        d.cf = P.d_cf;
        d.ci = P.d_ci;
        d.n_atom = T.n_atom;
        d.pos_out = T.nodes[0].out;
        d.pos_sens = T.nodes[0].sens;
        d.n_pot = T.n_pot;
        d.pot = T.pot;
        d.sens_zero_cnt = (int)T.sens_zero.size();
        d.sens_zero = T.szero;
        d.mom = T.mom;
        d.error = T.err;
        d.n_lanes = P.n_lanes;
        d.bar = T.bar;
        d.tred = T.tred;
        d.seed = base_seed + (uint32_t)k;
        // OrnsteinUhlenbeckThermostat parameters (src/thermostat.h:9-12, src/main.cpp:775-781)
        float delta_t = (float)thermostat_interval * 3.f * dt;
        float ms = std::exp(-delta_t / thermostat_timescale);
        d.mom_scale = ms;
        d.noise_scale = std::sqrt(s.temperature * (1.f - ms * ms));
        d.n_invocations = 0;
        d.temp0 = s.temp0;
        d.prof = (profile && k < (int)prof.size()) ? prof[k] : nullptr;
        return d;
    }

    void create(const vector<string>& paths, const vector<float>& temps, int n_atom_hint) {
        std::map<string, std::shared_ptr<Program>> by_path;
        // UPSIDE_CAPACITY_SCALE (e.g. 0.05) starts with small neighbour-list
        // capacities; overflow regrowth then doubles them as needed
        if (const char* e = std::getenv("UPSIDE_CAPACITY_SCALE")) {
            double v = std::atof(e);
            if (v > 0.0 && v <= 64.0) cap_scale = v;
        }
        hip_check(hipStreamCreate(&stream), "stream");
        hip_check(hipEventCreate(&ev0), "event");
        hip_check(hipEventCreate(&ev1), "event");
        for (size_t k = 0; k < paths.size(); ++k) {
            auto it = by_path.find(paths[k]);
            if (it == by_path.end()) {
                auto P = std::make_shared<Program>();
                P->path = paths[k];
                P->load(n_atom_hint, cap_scale);
                progs.push_back(P);
                it = by_path.emplace(paths[k], P).first;
            }
            System s;
            s.prog = it->second;
            s.temperature = k < temps.size() ? temps[k] : 1.f;
            s.temp0 = s.temperature;
            sys.push_back(s);
        }
        for (auto& s : sys) alloc_system(s);
        desc.clear();
        for (size_t k = 0; k < sys.size(); ++k) desc.push_back(make_desc((int)k));
        hip_check(hipMalloc((void**)&d_desc, desc.size() * sizeof(SysDesc)), "descriptors");
        push_desc();
        if (const char* e = std::getenv("UPSIDE_TEAM")) {  // fixed team size (tests, experiments)
            int t = std::atoi(e);
            if (t >= 1) team = std::min(t, max_team());
        }
    }

    void push_desc() {
        hip_check(hipMemcpy(d_desc, desc.data(), desc.size() * sizeof(SysDesc), hipMemcpyHostToDevice), "descriptors");
    }
    void pull_desc() {
        hip_check(hipMemcpy(desc.data(), d_desc, desc.size() * sizeof(SysDesc), hipMemcpyDeviceToHost), "descriptors");
    }

    bool overflowed() {
        bool of = false;
        for (auto& s : sys) {
            const Topology& T = s.prog->T;
            for (auto& G : T.graphs) {
                int st[4];
                hip_check(hipMemcpy(st, s.iarena + G.state, sizeof(st), hipMemcpyDeviceToHost), "graph state");
                if (st[1]) of = true;
            }
        }
        return of;
    }

    // Re-lower every program with larger list capacities, keeping positions,
    // momenta and thermostat counters.
    void grow() {
        if (cap_scale * 2.0 > 128.0) fail("neighbour list capacity growth limit reached");
        relower(cap_scale * 2.0);
    }

    // Lower every program again with list capacities scaled by `scale` (larger
    // or smaller) and reallocate the systems, keeping positions, momenta and
    // thermostat counters.
    void relower(double scale) {
        cap_scale = scale;
        vector<vector<float>> pos(sys.size()), mom(sys.size());
        pull_desc();
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            pos[k].resize(T.n_atom * 3);
            mom[k].resize(T.n_atom * 3);
            hip_check(hipMemcpy(pos[k].data(), sys[k].arena + T.nodes[0].out, pos[k].size() * 4, hipMemcpyDeviceToHost), "pos");
            hip_check(hipMemcpy(mom[k].data(), sys[k].arena + T.mom, mom[k].size() * 4, hipMemcpyDeviceToHost), "mom");
        }
        for (auto& P : progs) P->load(P->T.n_atom, cap_scale);
        for (size_t k = 0; k < sys.size(); ++k) {
            sys[k].release();
            alloc_system(sys[k]);
            uint64_t inv = desc[k].n_invocations;
            desc[k] = make_desc((int)k);
            desc[k].n_invocations = inv;
            const Topology& T = sys[k].prog->T;
            hip_check(hipMemcpy(sys[k].arena + T.nodes[0].out, pos[k].data(), pos[k].size() * 4, hipMemcpyHostToDevice), "pos");
            hip_check(hipMemcpy(sys[k].arena + T.mom, mom[k].data(), mom[k].size() * 4, hipMemcpyHostToDevice), "mom");
        }
        push_desc();
    }

    int lane_split = 2;  // sections of four-lane programs: 4 quarters or 2 halves (RunArgs::lane_split)
    int team = 1;   // workgroups per system (set_team / autotune); with per-system teams their largest
    // per-system team sizes (set_teams; empty: every system has `team`) and the
    // device block map of their launches (RunArgs::blockmap)
    vector<int> teams;
    int* d_blockmap = nullptr;
    int n_blocks = 0, map_cap = 0;
    int team_of(size_t s) const { return teams.empty() ? team : teams[s]; }
    long long team_total() const {
        long long t = 0;
        for (size_t s = 0; s < sys.size(); ++s) t += team_of(s);
        return t;
    }
    int bp_layout = 0;  // where rotamer solves run (RunArgs::bp_layout)
    // rotamer solve schedule (RunArgs::bp_mode and its parameters) and the
    // warm-start network's device copy (schedule 2)
    int bp_mode = 0;
    float bp_residual = 1e-5f;
    int bp_budget = 200;
    int bp_klocal = 4;
    std::unique_ptr<MlpParams> mlp;
    // replica exchange inside the next run's launch (interval 0: none); the
    // owner restores the exchange state before every launch attempt of the
    // run (a retry after a list overflow repeats the attempts)
    RexDev rex_dev{};
    std::function<void()> rex_restore;
    bool team_local = false;  // the last team launch ran every team on one XCD

    // Largest team every system of the batch can have with all workgroups
    // resident at once (cooperative launch), capped at MAX_TEAM.
    int max_team() const {
        int blocks = upside_max_resident_blocks();
        int groups = ((int)sys.size() + N_XCD - 1) / N_XCD;  // team launches pad to N_XCD systems
        int t = blocks / (N_XCD * std::max(groups, 1));
        return std::max(1, std::min(t, MAX_TEAM));
    }

    void reset_barriers() {  // the barrier counters must be multiples of a new team size
        for (auto& sy : sys) hip_check(hipMemset(sy.iarena + sy.prog->T.bar, 0, BAR_WORDS * sizeof(int)), "barrier reset");
    }

    void set_team(int t) {
        if (t < 1 || t > max_team()) fail("team size " + std::to_string(t) + " not possible (max " +
                                          std::to_string(max_team()) + " for this batch)");
        if (t != team || !teams.empty()) reset_barriers();
        team = t;
        teams.clear();
    }

    bool pack_teams(const vector<int>& t, vector<int>* map) const {
        return plan_team_map(t, upside_max_resident_blocks() / N_XCD, map);
    }

    void set_teams(const vector<int>& v) {
        if (v.size() != sys.size()) fail("set_teams: one team size per system expected");
        if (sys.size() >= (1u << 17)) fail("set_teams: too many systems for per-system teams");
        for (int g : v)
            if (g < 1 || g > MAX_TEAM) fail("set_teams: team size " + std::to_string(g) + " out of range");
        vector<int> map;
        if (!pack_teams(v, &map))
            fail("set_teams: the teams do not fit on the device at once (" +
                 std::to_string(upside_max_resident_blocks() / N_XCD) + " workgroups per XCD)");
        if ((int)map.size() > map_cap) {
            (void)hipFree(d_blockmap);
            d_blockmap = nullptr;
            map_cap = 0;
            hip_check(hipMalloc(&d_blockmap, map.size() * sizeof(int)), "block map");
            map_cap = (int)map.size();
        }
        hip_check(hipMemcpy(d_blockmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice), "block map");
        n_blocks = (int)map.size();
        if (v != teams) reset_barriers();
        teams = v;
        team = *std::max_element(v.begin(), v.end());
    }

    // Balanced per-system teams (plan_balanced_teams) within budget
    // workgroups (<= 0: the device's resident workgroups).
    vector<int> balanced_teams(int budget) const {
        const int resident = upside_max_resident_blocks();
        if (budget <= 0) budget = resident;
        vector<int> atoms(sys.size());
        for (size_t s = 0; s < sys.size(); ++s) atoms[s] = sys[s].prog->T.n_atom;
        vector<int> t = plan_balanced_teams(atoms, budget, resident / N_XCD);
        if (t.empty()) fail("balanced teams: more systems than resident workgroups");
        return t;
    }

    // Pick the team size and rotamer-solve layout with the shortest run of
    // n_round rounds: every candidate pair starts from the same snapshot
    // (positions, momenta, thermostat counters), which is restored
    // afterwards, so tuning leaves the state as it was.
    // ms_out[k * n_lay + l] receives the kernel time of team cand[k] with
    // layout lay[l] (-1 if skipped).  lay == nullptr tunes the team alone.
    void autotune(int n_round, const int* cand, int n_cand, float* ms_out, const int* lay = nullptr,
                  int n_lay = 0) {
        const int cur_layout = bp_layout;
        if (!lay) { lay = &cur_layout; n_lay = 1; }
        for (int l = 0; l < n_lay; ++l)
            if (lay[l] < 0 || lay[l] > 3) fail("rotamer solve layout must be 0, 1, 2 or 3");
        vector<vector<float>> pos(sys.size()), mom(sys.size());
        pull_desc();
        vector<SysDesc> saved = desc;
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            pos[k].resize(T.n_atom * 3);
            mom[k].resize(T.n_atom * 3);
            hip_check(hipMemcpy(pos[k].data(), sys[k].arena + T.nodes[0].out, pos[k].size() * 4, hipMemcpyDeviceToHost), "pos");
            hip_check(hipMemcpy(mom[k].data(), sys[k].arena + T.mom, mom[k].size() * 4, hipMemcpyDeviceToHost), "mom");
        }
        auto restore = [&]() {
            for (size_t k = 0; k < sys.size(); ++k) {
                const Topology& T = sys[k].prog->T;
                hip_check(hipMemcpy(sys[k].arena + T.nodes[0].out, pos[k].data(), pos[k].size() * 4, hipMemcpyHostToDevice), "pos");
                hip_check(hipMemcpy(sys[k].arena + T.mom, mom[k].data(), mom[k].size() * 4, hipMemcpyHostToDevice), "mom");
                desc[k].n_invocations = saved[k].n_invocations;
            }
            push_desc();
        };
        int best = teams.empty() ? team : 0, best_lay = bp_layout;
        float best_ms = 3.0e38f;
        int limit = max_team();
        vector<int> balanced;
        for (int k = 0; k < n_cand; ++k)
            if (cand[k] == 0 && balanced.empty()) balanced = balanced_teams(0);
        for (int k = 0; k < n_cand; ++k)
            for (int l = 0; l < n_lay; ++l) {
                if (ms_out) ms_out[k * n_lay + l] = -1.f;
                if (cand[k] < 0 || cand[k] > limit) continue;
                if (cand[k] == 1 && l > 0) continue;  // a team of one solves on its workgroup anyway
                if (cand[k] == 0) set_teams(balanced);
                else set_team(cand[k]);
                bp_layout = lay[l];
                run(n_round, 0, nullptr, 0, (int)round_num);
                if (ms_out) ms_out[k * n_lay + l] = last_ms;
                if (last_ms < best_ms) { best_ms = last_ms; best = cand[k]; best_lay = lay[l]; }
                restore();
            }
        if (best == 0) set_teams(balanced);
        else set_team(best);
        bp_layout = best_lay;
        // then the section split of four-lane programs (quarters or halves;
        // the same results either way) where a team has eight workgroups
        bool four = false;
        for (size_t s = 0; s < sys.size(); ++s) four = four || (sys[s].prog->n_lanes == 4 && team_of(s) >= 8);
        if (four) {
            int best_split = lane_split;
            float split_ms = 3.0e38f;
            for (int split : {2, 4}) {
                lane_split = split;
                run(n_round, 0, nullptr, 0, (int)round_num);
                if (last_ms < split_ms) { split_ms = last_ms; best_split = split; }
                restore();
            }
            lane_split = best_split;
        }
    }

    void launch(const RunArgs& a0) {
        RunArgs a = a0;
        a.team = team;
        a.lane_split = lane_split;
        if (!teams.empty()) {
            a.blockmap = d_blockmap;
            a.n_blocks = n_blocks;
        }
        a.bp_layout = bp_layout;
        a.bp_mode = bp_mode;
        a.bp_residual = bp_residual;
        a.bp_budget = bp_budget;
        a.bp_klocal = bp_klocal;
        if (mlp) a.mlp = mlp->dev;
        hip_check(hipEventRecord(ev0, stream), "event");
        hip_check(upside_launch_engine(d_desc, (int)sys.size(), a, stream), "engine launch");
        hip_check(hipEventRecord(ev1, stream), "event");
        hip_check(hipStreamSynchronize(stream), "engine run");
        (void)hipEventElapsedTime(&last_ms, ev0, ev1);
        if (team > 1) team_local = true;
        if (team > 1)
            for (size_t s = 0; s < sys.size(); ++s) {
                const System& sy = sys[s];
                if (team_of(s) < 2) continue;
                unsigned w[BAR_WORDS];
                hip_check(hipMemcpy(w, sy.iarena + sy.prog->T.bar, sizeof(w), hipMemcpyDeviceToHost), "barrier");
                auto word64 = [&](int k) { return (unsigned long long)w[k] | ((unsigned long long)w[k + 1] << 32); };
                const unsigned long long g = (unsigned long long)team_of(s), lo = g / 2, hi = g - g / 2;
                if (w[2]) fail("team barrier wait timed out: results of this run are void");
                bool quarters_ok = true;
                if (sy.prog->n_lanes == 4 && g >= 8)
                    for (unsigned long long j = 0; j < 4; ++j)
                        quarters_ok = quarters_ok && word64(BAR_Q + 2 * (int)j) % ((j + 1) * g / 4 - j * g / 4) == 0;
                if (word64(0) % g || (g > 2 && word64(BAR_SUB1) % (g - 1)) || (g >= 4 && word64(BAR_LO) % lo) ||
                    (g >= 4 && word64(BAR_HI) % hi) || !quarters_ok)
                    fail("team barrier arrivals unbalanced at the end of a launch: results of this run are void");
                team_local = team_local && w[3] != 0;
            }
    }

    // Evaluate potential and derivative of every system at its positions.
    void evaluate(int mode) {
        RunArgs a;
        std::memset(&a, 0, sizeof(a));
        a.mode = mode;
        for (;;) {  // grow() fails once a capacity reaches its limit
            launch(a);
            if (!overflowed()) break;
            grow();
        }
    }

    void run(int n_round, int frame_interval, float* frame_pot, int n_frames, int first_round) {
        // snapshot for a retry after list overflow
        vector<vector<float>> pos(sys.size()), mom(sys.size());
        pull_desc();
        vector<SysDesc> saved = desc;
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            pos[k].resize(T.n_atom * 3);
            mom[k].resize(T.n_atom * 3);
            hip_check(hipMemcpy(pos[k].data(), sys[k].arena + T.nodes[0].out, pos[k].size() * 4, hipMemcpyDeviceToHost), "pos");
            hip_check(hipMemcpy(mom[k].data(), sys[k].arena + T.mom, mom[k].size() * 4, hipMemcpyDeviceToHost), "mom");
        }
        float* d_frames = nullptr;
        if (frame_pot && n_frames > 0)
            hip_check(hipMalloc((void**)&d_frames, sys.size() * n_frames * sizeof(float)), "frames");
        RunArgs a;
        std::memset(&a, 0, sizeof(a));
        a.mode = 0;
        a.n_round = n_round;
        a.dt = dt;
        a.thermostat_interval = thermostat_interval;
        a.first_round = first_round;
        a.frame_interval = frame_interval;
        a.frame_pot = d_frames;
        a.n_frames = n_frames;
        a.anneal = anneal;
        a.anneal_factor = anneal_factor;
        a.anneal_start = anneal_start;
        a.anneal_end = anneal_end;
        a.anneal_duration = anneal_duration;
        a.rex = rex_dev;
        a.rex.n_wg = (int)team_total();  // the workgroups that meet at its grid barriers
        for (;;) {
            if (a.rex.interval > 0) {
                rex_restore();
                hip_check(hipMemset(a.rex.grid_bar, 0, 2 * sizeof(unsigned long long)), "grid barrier");
            }
            launch(a);
            if (a.rex.interval > 0) {
                unsigned flag = 0;
                hip_check(hipMemcpy(&flag, a.rex.grid_flag, sizeof(flag), hipMemcpyDeviceToHost), "grid barrier");
                if (flag && !overflowed()) fail("replica exchange grid barrier timed out: results of this run are void");
            }
            if (!overflowed()) break;
            grow();
            for (size_t k = 0; k < sys.size(); ++k) {
                const Topology& T = sys[k].prog->T;
                hip_check(hipMemcpy(sys[k].arena + T.nodes[0].out, pos[k].data(), pos[k].size() * 4, hipMemcpyHostToDevice), "pos");
                hip_check(hipMemcpy(sys[k].arena + T.mom, mom[k].data(), mom[k].size() * 4, hipMemcpyHostToDevice), "mom");
                desc[k].n_invocations = saved[k].n_invocations;
            }
            push_desc();
        }
        if (d_frames) {
            hip_check(hipMemcpy(frame_pot, d_frames, sys.size() * n_frames * sizeof(float), hipMemcpyDeviceToHost), "frames");
            (void)hipFree(d_frames);
        }
        pull_desc();
    }

    // Checkpoint of the whole device state: a header, then per system its
    // thermostat counter and temperature and its float and int arenas
    // (positions, momenta, Verlet caches, pair lists, BP state).  Loading it
    // into a batch of the same .up files resumes bit for bit: the pair-list
    // row order depends on when each cache was last rebuilt, so positions and
    // momenta alone would resume only to rounding.
    struct StateHeader {
        char magic[8];
        int version, n_sys, thermostat_interval;
        uint32_t base_seed;
        double cap_scale;
        uint64_t round_num;
        float dt, timescale;
    };
    struct SysState {
        long long nf, ni;
        uint64_t n_invocations;
        float temperature, mom_scale, noise_scale;
        uint32_t seed;
        float temp0, pad;
    };
    size_t state_size() const {
        size_t n = sizeof(StateHeader);
        for (auto& sy : sys)
            n += sizeof(SysState) + (size_t)std::max(sy.prog->T.nf, 4) * 4 + (size_t)std::max(sy.prog->T.ni, 4) * 4;
        return n;
    }
    void save_state(char* out, size_t size) {
        if (size < state_size()) fail("save_state: buffer of " + std::to_string(size) + " bytes, need " +
                                      std::to_string(state_size()));
        hip_check(hipStreamSynchronize(stream), "sync");
        pull_desc();
        StateHeader h;
        std::memset(&h, 0, sizeof(h));
        std::memcpy(h.magic, "UPSTATE1", 8);
        h.version = 1;
        h.n_sys = (int)sys.size();
        h.thermostat_interval = thermostat_interval;
        h.base_seed = base_seed;
        h.cap_scale = cap_scale;
        h.round_num = round_num;
        h.dt = dt;
        h.timescale = thermostat_timescale;
        std::memcpy(out, &h, sizeof(h));
        out += sizeof(h);
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            SysState st;
            std::memset(&st, 0, sizeof(st));
            st.nf = std::max(T.nf, 4);
            st.ni = std::max(T.ni, 4);
            st.n_invocations = desc[k].n_invocations;
            st.temperature = sys[k].temperature;
            st.mom_scale = desc[k].mom_scale;
            st.noise_scale = desc[k].noise_scale;
            st.seed = desc[k].seed;
            st.temp0 = sys[k].temp0;
            std::memcpy(out, &st, sizeof(st));
            out += sizeof(st);
            hip_check(hipMemcpy(out, sys[k].arena, st.nf * 4, hipMemcpyDeviceToHost), "save arena");
            out += st.nf * 4;
            hip_check(hipMemcpy(out, sys[k].iarena, st.ni * 4, hipMemcpyDeviceToHost), "save int arena");
            out += st.ni * 4;
        }
    }
    void load_state(const char* in, size_t size) {
        StateHeader h;
        if (size < sizeof(h)) fail("load_state: truncated state");
        std::memcpy(&h, in, sizeof(h));
        if (std::memcmp(h.magic, "UPSTATE1", 8) || h.version != 1) fail("load_state: not a saved batch state");
        if (h.n_sys != (int)sys.size())
            fail("load_state: state of " + std::to_string(h.n_sys) + " systems, batch of " +
                 std::to_string(sys.size()));
        if (!(h.cap_scale > 0.0 && h.cap_scale <= 128.0)) fail("load_state: list capacity scale out of range");
        hip_check(hipStreamSynchronize(stream), "sync");
        // validate everything before touching the batch: the arena sizes the
        // programs have when lowered at the saved capacity scale
        std::map<const Program*, std::array<int, 2>> sizes;
        for (auto& P : progs) {
            if (h.cap_scale == cap_scale) {
                sizes[P.get()] = {std::max(P->T.nf, 4), std::max(P->T.ni, 4)};
            } else {
                Topology t;
                t.cap_scale = h.cap_scale;
                H5 h5(P->path);
                load_topology(t, h5, P->T.n_atom);
                sizes[P.get()] = {std::max(t.nf, 4), std::max(t.ni, 4)};
            }
        }
        vector<SysState> st(sys.size());
        const char* p = in + sizeof(h);
        size_t need = sizeof(h);
        for (size_t k = 0; k < sys.size(); ++k) {
            if (size < need + sizeof(SysState)) fail("load_state: truncated state");
            std::memcpy(&st[k], p, sizeof(SysState));
            const auto& z = sizes[sys[k].prog.get()];
            if (st[k].nf != z[0] || st[k].ni != z[1])
                fail("load_state: system " + std::to_string(k) + " was saved from a different .up file");
            need += sizeof(SysState) + (size_t)st[k].nf * 4 + (size_t)st[k].ni * 4;
            p += sizeof(SysState) + (size_t)st[k].nf * 4 + (size_t)st[k].ni * 4;
        }
        if (size < need) fail("load_state: truncated state");
        if (h.cap_scale != cap_scale) relower(h.cap_scale);  // larger or smaller, as saved
        p = in + sizeof(h);
        pull_desc();
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            p += sizeof(SysState);
            hip_check(hipMemcpy(sys[k].arena, p, st[k].nf * 4, hipMemcpyHostToDevice), "load arena");
            p += st[k].nf * 4;
            hip_check(hipMemcpy(sys[k].iarena, p, st[k].ni * 4, hipMemcpyHostToDevice), "load int arena");
            p += st[k].ni * 4;
            // barrier counters restart at zero (a multiple of any team size)
            hip_check(hipMemset(sys[k].iarena + T.bar, 0, BAR_WORDS * sizeof(int)), "barrier reset");
            desc[k].n_invocations = st[k].n_invocations;
            desc[k].mom_scale = st[k].mom_scale;
            desc[k].noise_scale = st[k].noise_scale;
            desc[k].seed = st[k].seed;
            sys[k].temp0 = st[k].temp0;
            desc[k].temp0 = st[k].temp0;
            sys[k].temperature = st[k].temperature;
        }
        dt = h.dt;
        thermostat_interval = h.thermostat_interval;
        base_seed = h.base_seed;
        round_num = h.round_num;
        thermostat_timescale = h.timescale;
        have_mom = true;
        push_desc();
    }

    // Fresh-start thermalisation (src/main.cpp:800-803): p = 0 then one
    // thermostat application with the run's parameters.
    void thermalize() {
        for (size_t k = 0; k < sys.size(); ++k) {
            const Topology& T = sys[k].prog->T;
            hip_check(hipMemset(sys[k].arena + T.mom, 0, T.n_atom * 3 * sizeof(float)), "mom");
        }
        RunArgs a;
        std::memset(&a, 0, sizeof(a));
        a.mode = 0;
        a.n_round = -1;  // thermostat only
        a.thermostat_interval = 1;
        launch(a);
        pull_desc();
        have_mom = true;
    }

    float potential(int k) {
        const Topology& T = sys[k].prog->T;
        vector<float> p(std::max(T.n_pot, 1));
        hip_check(hipMemcpy(p.data(), sys[k].arena + T.pot, p.size() * 4, hipMemcpyDeviceToHost), "potential");
        float e = 0.f;
        for (int i = 0; i < T.n_pot; ++i) e += p[i];
        return e;
    }

    const Node* node(int k, const char* name) {
        for (auto& n : sys[k].prog->T.nodes)
            if (n.name == name) return &n;
        fail(string("name not found: ") + name);
    }

    void copy_pos_in(int k, const float* pos) {
        const Topology& T = sys[k].prog->T;
        hip_check(hipMemcpy(sys[k].arena + T.nodes[0].out, pos, T.n_atom * 3 * sizeof(float), hipMemcpyHostToDevice), "pos");
    }
};

static void set_initial_positions(Batch& B) {
    for (size_t k = 0; k < B.sys.size(); ++k) {
        H5 h5(B.sys[k].prog->path);
        auto p = h5.rf("/input/pos");
        const Topology& T = B.sys[k].prog->T;
        if ((int)p.size() < T.n_atom * 3) fail("/input/pos smaller than n_atom");
        B.copy_pos_in((int)k, p.data());
    }
}

}  // namespace upside

// ================================================================ C ABI ===
using namespace upside;

struct DerivEngine {
    Batch batch;
    bool evaluated = false;   // an evaluation has set every node's outputs and sensitivities
};
// /output logger of one system: the basic-level datasets of the reference's
// simulation loop (src/main.cpp:807-831 pos, mom, kinetic, potential, time;
// :893 temperature) as extendable arrays with the first dimension unlimited
// and chunks of 100 frames (src/state_logger.h:21-39), appended one frame at a
// time and flushed on close.
// The BP state of the last evaluation on the host: node probabilities and
// marginals, BP edges with matrices and marginals, 1-rotamer pairs.  One
// fetch per array; a caller asking for several values of one evaluation
// passes the same RotFetch to every rotamer_values call.
struct RotFetch {
    bool filled = false;
    vector<float> prob, cur, off, P, M, p11;
    vector<int> bpf, bps;
    vector<std::array<int, 2>> e11;
    int nbp = 0;
};

static vector<float> rotamer_values(Batch& B, int sys, int op, const string& ln, RotFetch* cache = nullptr);

struct OutputLog {
    hid_t file = -1, group = -1;
    hid_t pos = -1, mom = -1, kinetic = -1, potential = -1, time = -1, temperature = -1;
    hid_t pivot_stats = -1;   // (n_frame, 2): accepted and attempted pivots since the last frame
    // detailed level (src/hbond.cpp:373-377, src/rotamer.cpp:708-710)
    hid_t hbond = -1, bad_solves = -1;
    hid_t replica_index = -1, replica_swaps = -1;   // src/main.cpp:210-232
    int n_partner = 0;
    int hb_node = -1, hb_n = 0, rot_stats = -1;
    // per-residue rotamer free energy, its parts and the 1-body energies
    // (src/rotamer.cpp:711-731), each (n_frame, n_node)
    int rot_op = -1, rot_nn = 0, rot_nprob = 0;
    vector<hid_t> rot_sets;
    vector<vector<float>> rot_rows;   // filled by the caller before frame()
    // per-node values of the detailed and extensive levels, one dataset each:
    // NL_RAMA rama_map_potential (src/rama_map_pot.cpp:49-55), NL_SIGMOID
    // sigmoid_coupling (src/environment.cpp:892-900), NL_OUTPUT the node's
    // output (environment_coverage, hbond_environment_coverage,
    // src/environment.cpp:304-310, :355-360), NL_VIRTUAL the first three
    // columns of infer_H_O's rows (virtual, src/hbond.cpp:53-58)
    enum { NL_RAMA, NL_SIGMOID, NL_OUTPUT, NL_VIRTUAL };
    struct NodeLog {
        hid_t d;
        int kind, node, n;
        vector<float> row;   // filled by the caller before frame()
    };
    vector<NodeLog> node_logs;
    long long n = 0;

    static hid_t earray(hid_t g, const char* name, hid_t type, const vector<hsize_t>& row) {
        vector<hsize_t> dims{0}, maxd{H5S_UNLIMITED}, chunk{100};
        for (hsize_t r : row) { dims.push_back(r); maxd.push_back(r); chunk.push_back(r); }
        hid_t sp = H5Screate_simple((int)dims.size(), dims.data(), maxd.data());
        hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(pl, (int)chunk.size(), chunk.data());
        hid_t d = H5Dcreate2(g, name, type, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
        H5Pclose(pl);
        H5Sclose(sp);
        if (d < 0) fail(string("output: cannot create dataset ") + name);
        return d;
    }
    static void append(hid_t d, hid_t type, const void* data, long long row) {
        hid_t sp = H5Dget_space(d);
        int rank = H5Sget_simple_extent_ndims(sp);
        vector<hsize_t> dims(rank);
        H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
        H5Sclose(sp);
        dims[0] = (hsize_t)row + 1;
        if (H5Dset_extent(d, dims.data()) < 0) fail("output: cannot extend dataset");
        sp = H5Dget_space(d);
        vector<hsize_t> start(rank, 0), count = dims;
        start[0] = (hsize_t)row;
        count[0] = 1;
        H5Sselect_hyperslab(sp, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr);
        hid_t ms = H5Screate_simple(rank, count.data(), nullptr);
        herr_t e = H5Dwrite(d, type, ms, sp, H5P_DEFAULT, data);
        H5Sclose(ms);
        H5Sclose(sp);
        if (e < 0) fail("output: write failed");
    }
    void open(const string& path, bool into_input, int n_atom, bool record_mom, bool mc, const Topology& T,
              int level, const vector<int>* partners) {
        if (into_input) {
            file = H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
            if (file >= 0 && H5Lexists(file, "output", H5P_DEFAULT) > 0) H5Ldelete(file, "output", H5P_DEFAULT);
        } else {
            file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        }
        if (file < 0) fail("output: cannot open " + path + " for writing");
        group = H5Gcreate2(file, "output", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (group < 0) fail("output: cannot create /output in " + path);
        const hsize_t na = (hsize_t)n_atom;
        pos = earray(group, "pos", H5T_NATIVE_FLOAT, {1, na, 3});
        if (record_mom) mom = earray(group, "mom", H5T_NATIVE_FLOAT, {1, na, 3});
        kinetic = earray(group, "kinetic", H5T_NATIVE_DOUBLE, {1});
        potential = earray(group, "potential", H5T_NATIVE_DOUBLE, {1});
        time = earray(group, "time", H5T_NATIVE_DOUBLE, {});
        temperature = earray(group, "temperature", H5T_NATIVE_DOUBLE, {1});
        if (mc) pivot_stats = earray(group, "pivot_stats", H5T_NATIVE_INT, {2});  // src/monte_carlo_sampler.h:32-36
        if (partners) {  // replica exchange: index, partners once, cumulative swaps per partner
            replica_index = earray(group, "replica_index", H5T_NATIVE_INT, {1});
            n_partner = (int)partners->size();
            if (n_partner > 0) {
                hid_t d = earray(group, "replica_swap_partner", H5T_NATIVE_INT, {});
                hsize_t dims[1] = {(hsize_t)n_partner};
                H5Dset_extent(d, dims);
                H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, partners->data());
                H5Dclose(d);
                replica_swaps = earray(group, "replica_cumulative_swaps", H5T_NATIVE_INT, {(hsize_t)n_partner, 2});
            }
        }
        if (level >= 1) {
            for (const Node& nd : T.nodes)
                if (nd.op >= 0 && T.ops[nd.op].kind == OP_PROTEIN_HBOND && hb_node < 0) {
                    hb_node = (int)(&nd - T.nodes.data());
                    hb_n = nd.n_elem;
                }
            if (hb_node >= 0) hbond = earray(group, "hbond", H5T_NATIVE_FLOAT, {(hsize_t)hb_n});
            if (!T.rots.empty()) {
                rot_stats = T.rots[0].stats;
                bad_solves = earray(group, "rotamer_bad_solves_cumulative", H5T_NATIVE_LLONG, {1});
                for (size_t o = 0; o < T.ops.size(); ++o)
                    if (T.ops[o].kind == OP_ROTAMER && T.ops[o].rot == 0) rot_op = (int)o;
                rot_nn = T.rots[0].n_node;
                rot_nprob = T.rots[0].n_prob;
                vector<string> names{"rotamer_free_energy", "rotamer_potential", "rotamer_entropy"};
                for (int q = 0; q < rot_nprob; ++q) names.push_back("rotamer_1body_energy" + std::to_string(q));
                if (rot_op >= 0)
                    for (auto& nm : names) rot_sets.push_back(earray(group, nm.c_str(), H5T_NATIVE_FLOAT, {(hsize_t)rot_nn}));
                rot_rows.resize(rot_sets.size());
            }
            // a dataset name is created once: the reference would fail on a
            // second node logging under the same name, here the first in
            // load order is logged
            std::set<string> made;
            auto add = [&](const char* name, int kind, int node, int n, vector<hsize_t> row) {
                if (made.count(name)) return;
                made.insert(name);
                node_logs.push_back({earray(group, name, H5T_NATIVE_FLOAT, row), kind, node, n, {}});
            };
            for (size_t k = 0; k < T.nodes.size(); ++k) {
                const Node& nd = T.nodes[k];
                if (nd.op < 0) continue;
                const OpRec& o = T.ops[nd.op];
                if (o.kind == OP_RAMA_MAP_POT && o.ia[5])
                    add("rama_map_potential", NL_RAMA, (int)k, o.ia[4], {(hsize_t)o.ia[4]});
                if (o.kind == OP_SIGMOID_COUPLING) add("sigmoid_coupling", NL_SIGMOID, (int)k, o.ia[0], {(hsize_t)o.ia[0]});
                if (level >= 2) {
                    if (o.kind == OP_ENV_COVERAGE)
                        add("environment_coverage", NL_OUTPUT, (int)k, nd.n_elem * nd.width, {(hsize_t)(nd.n_elem * nd.width)});
                    if (o.kind == OP_HB_ENV_COVERAGE)
                        add("hbond_environment_coverage", NL_OUTPUT, (int)k, nd.n_elem * nd.width,
                            {(hsize_t)(nd.n_elem * nd.width)});
                    if (o.kind == OP_INFER_HO) add("virtual", NL_VIRTUAL, (int)k, nd.n_elem * 3, {(hsize_t)nd.n_elem, 3});
                }
            }
        }
        // the reference records its command line here (src/main.cpp:734-736)
        const char* inv = "libupside batch (upside_batch_run_logged)";
        hid_t st = H5Tcopy(H5T_C_S1);
        H5Tset_size(st, strlen(inv));
        hid_t sp = H5Screate(H5S_SCALAR);
        hid_t a = H5Acreate2(group, "invocation", st, sp, H5P_DEFAULT, H5P_DEFAULT);
        if (a >= 0) { H5Awrite(a, st, inv); H5Aclose(a); }
        H5Sclose(sp);
        H5Tclose(st);
    }
    vector<float> hb_values;   // filled by the caller before frame() when hbond >= 0
    long long n_bad = 0;
    int rep_index = 0;
    vector<int> rep_swaps;     // [n_partner][2] successes, attempts
    void frame(const float* p, const float* m, int n_atom, double kin, double pot, double t, double temp,
               const int* mc_stats) {
        append(pos, H5T_NATIVE_FLOAT, p, n);
        if (mom >= 0) append(mom, H5T_NATIVE_FLOAT, m, n);
        append(kinetic, H5T_NATIVE_DOUBLE, &kin, n);
        append(potential, H5T_NATIVE_DOUBLE, &pot, n);
        append(time, H5T_NATIVE_DOUBLE, &t, n);
        append(temperature, H5T_NATIVE_DOUBLE, &temp, n);
        if (pivot_stats >= 0) append(pivot_stats, H5T_NATIVE_INT, mc_stats, n);
        if (hbond >= 0) append(hbond, H5T_NATIVE_FLOAT, hb_values.data(), n);
        if (bad_solves >= 0) append(bad_solves, H5T_NATIVE_LLONG, &n_bad, n);
        for (size_t q = 0; q < rot_sets.size(); ++q) append(rot_sets[q], H5T_NATIVE_FLOAT, rot_rows[q].data(), n);
        for (auto& nl : node_logs) append(nl.d, H5T_NATIVE_FLOAT, nl.row.data(), n);
        if (replica_index >= 0) append(replica_index, H5T_NATIVE_INT, &rep_index, n);
        if (replica_swaps >= 0) append(replica_swaps, H5T_NATIVE_INT, rep_swaps.data(), n);
        ++n;
        (void)n_atom;
    }
    void close() {
        for (hid_t* d : {&pos, &mom, &kinetic, &potential, &time, &temperature, &pivot_stats, &hbond, &bad_solves,
                         &replica_index, &replica_swaps})
            if (*d >= 0) { H5Dclose(*d); *d = -1; }
        for (hid_t& d : rot_sets)
            if (d >= 0) { H5Dclose(d); d = -1; }
        if (group >= 0) { H5Gclose(group); group = -1; }
        if (file >= 0) { H5Fflush(file, H5F_SCOPE_LOCAL); H5Fclose(file); file = -1; }
    }
    ~OutputLog() { close(); }
};

struct UpsideBatch {
    Batch batch;
    std::vector<std::unique_ptr<OutputLog>> out;
    bool record_mom = false;
    int mc_interval = 0;                    // rounds between pivot moves (0: none)
    int recenter = 0;                       // before each frame: 0 none, 1 xyz, 2 xy (src/main.cpp:576-577, :947)
    // replica exchange inside run_logged (ReplicaExchange, src/main.cpp:159-316)
    int rex_interval = 0, rex_criterion = 0;
    vector<vector<std::array<int, 2>>> rex_sets;        // swap sets, pairs in order
    vector<vector<std::array<long long, 2>>> rex_stats; // per pair: successes, attempts
    vector<vector<std::array<int, 2>>> rex_partner;     // per system: (set, pair) it takes part in
    vector<int> replica_of;                            // replica index held by each system
    long long rex_attempts = 0;
    uint64_t rex_next = 0;
    static uint64_t next_swap_round(uint64_t after, int ri) {
        // a segment of the reference loop ends at the first multiple of the
        // interval at least two rounds past its start (src/main.cpp:918-990)
        uint64_t e = after + 2;
        return (e + ri - 1) / ri * ri;
    }
    // device copies of the exchange configuration and state for runs that
    // exchange inside the launch (one allocation, rebuilt with the sets)
    char* rex_mem = nullptr;
    size_t rex_mem_size = 0;
    RexDev rex_dev{};
    ~UpsideBatch() { (void)hipFree(rex_mem); }

    // The exchange can run inside the launch when every system evaluates the
    // same potential (one lowered program), temperatures are fixed (no
    // annealing) and every workgroup of the launch can be resident at once.
    bool device_rex_ok() const {
        if (rex_interval <= 0 || batch.anneal) return false;
        for (auto& sy : batch.sys)
            if (sy.prog != batch.sys[0].prog) return false;
        return batch.team_total() <= (long long)upside_max_resident_blocks();
    }

    void rex_alloc() {
        const int ns = (int)batch.sys.size();
        int np = 0;
        for (auto& st : rex_sets) np += (int)st.size();
        const int stride = (batch.sys[0].prog->T.n_atom * 3 + 63) / 64 * 64;
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t sizes[] = {(rex_sets.size() + 1) * 4, (size_t)2 * np * 4, 2 * 8, (size_t)2 * np * 8,
                                (size_t)ns * 4, (size_t)ns * 4, (size_t)ns * 4, (size_t)ns * stride * 4,
                                (size_t)ns * 4, 4 * 8};
        size_t total = 0;
        for (size_t b : sizes) total += up(b);
        if (total != rex_mem_size) {
            (void)hipFree(rex_mem);
            rex_mem = nullptr;
            hip_check(hipMalloc((void**)&rex_mem, total), "replica exchange state");
            rex_mem_size = total;
        }
        char* p = rex_mem;
        auto take = [&](size_t b) { char* q = p; p += up(b); return q; };
        rex_dev.set_off = (const int*)take(sizes[0]);
        rex_dev.pairs = (const int*)take(sizes[1]);
        rex_dev.state = (long long*)take(sizes[2]);
        rex_dev.stats = (long long*)take(sizes[3]);
        rex_dev.replica_of = (int*)take(sizes[4]);
        rex_dev.beta = (const float*)take(sizes[5]);
        rex_dev.energy = (float*)take(sizes[6]);
        rex_dev.stage = (float*)take(sizes[7]);
        rex_dev.partner = (int*)take(sizes[8]);
        char* g = take(sizes[9]);
        rex_dev.grid_bar = (unsigned long long*)g;
        rex_dev.grid_flag = (unsigned*)(g + 16);
        rex_dev.stage_stride = stride;
        rex_dev.interval = rex_interval;
        rex_dev.criterion = rex_criterion;
        rex_dev.n_sets = (int)rex_sets.size();
        rex_dev.seed = batch.base_seed;
        vector<int> off{0}, pr;
        for (auto& st : rex_sets) {
            for (auto& q : st) { pr.push_back(q[0]); pr.push_back(q[1]); }
            off.push_back((int)pr.size() / 2);
        }
        hip_check(hipMemcpy((void*)rex_dev.set_off, off.data(), off.size() * 4, hipMemcpyHostToDevice), "rex");
        if (!pr.empty()) hip_check(hipMemcpy((void*)rex_dev.pairs, pr.data(), pr.size() * 4, hipMemcpyHostToDevice), "rex");
    }

    // host state -> device (before every launch attempt of a run)
    void rex_upload() {
        const int ns = (int)batch.sys.size();
        long long st[2] = {rex_attempts, (long long)rex_next};
        vector<long long> stats;
        for (auto& v : rex_stats)
            for (auto& q : v) { stats.push_back(q[0]); stats.push_back(q[1]); }
        vector<float> beta(ns);
        for (int k = 0; k < ns; ++k) beta[k] = 1.f / batch.sys[k].temperature;
        hip_check(hipMemcpy(rex_dev.state, st, sizeof(st), hipMemcpyHostToDevice), "rex");
        if (!stats.empty()) hip_check(hipMemcpy(rex_dev.stats, stats.data(), stats.size() * 8, hipMemcpyHostToDevice), "rex");
        hip_check(hipMemcpy(rex_dev.replica_of, replica_of.data(), ns * 4, hipMemcpyHostToDevice), "rex");
        hip_check(hipMemcpy((void*)rex_dev.beta, beta.data(), ns * 4, hipMemcpyHostToDevice), "rex");
        hip_check(hipMemset(rex_dev.grid_flag, 0, sizeof(unsigned)), "rex");
    }

    // device state -> host (after a run)
    void rex_download() {
        const int ns = (int)batch.sys.size();
        long long st[2];
        hip_check(hipMemcpy(st, rex_dev.state, sizeof(st), hipMemcpyDeviceToHost), "rex");
        rex_attempts = st[0];
        rex_next = (uint64_t)st[1];
        size_t np = 0;
        for (auto& v : rex_stats) np += v.size();
        vector<long long> stats(2 * np);
        if (np) hip_check(hipMemcpy(stats.data(), rex_dev.stats, stats.size() * 8, hipMemcpyDeviceToHost), "rex");
        size_t k = 0;
        for (auto& v : rex_stats)
            for (auto& q : v) { q[0] = stats[k]; q[1] = stats[k + 1]; k += 2; }
        hip_check(hipMemcpy(replica_of.data(), rex_dev.replica_of, ns * 4, hipMemcpyDeviceToHost), "rex");
    }

    vector<std::array<long long, 2>> mc_total;  // accepted, attempted since set_monte_carlo
    vector<std::array<int, 2>> mc_frame;        // the same since the last logged frame

    // One pivot move per system at round `round` (MonteCarloSampler::
    // monte_carlo_step, src/monte_carlo_sampler.cpp:255-283): potentials of
    // every system before and after the proposals are evaluated together on
    // the GPU; generator (seed, stream 2, 0, round), first draw for the
    // proposal, second for the Metropolis test when it is needed.
    void mc_step(uint64_t round) {
        const int ns = (int)batch.sys.size();
        batch.pull_desc();
        batch.evaluate(1);
        vector<float> e_old(ns), dl(ns);
        vector<vector<float>> pos(ns), keep(ns);
        for (int k = 0; k < ns; ++k) e_old[k] = batch.potential(k);
        for (int k = 0; k < ns; ++k) {
            const Topology& T = batch.sys[k].prog->T;
            pos[k].resize(T.n_atom * 3);
            hip_check(hipMemcpy(pos[k].data(), batch.sys[k].arena + T.nodes[0].out, pos[k].size() * 4,
                                hipMemcpyDeviceToHost), "mc pos");
            keep[k] = pos[k];
            uint32_t key[4] = {batch.desc[k].seed, 2u, 0u, 0u};
            uint32_t ctr[4] = {(uint32_t)(round & 0xffffffffull), (uint32_t)(round >> 32), 0u, 0u};
            uint32_t bits[4];
            threefry_host(ctr, key, bits);
            float u[4] = {u01_open_closed(bits[0]), u01_open_closed(bits[1]), u01_open_closed(bits[2]),
                          u01_open_closed(bits[3])};
            dl[k] = batch.sys[k].prog->pivots.propose(pos[k].data(), u);
            hip_check(hipMemcpy(batch.sys[k].arena + T.nodes[0].out, pos[k].data(), pos[k].size() * 4,
                                hipMemcpyHostToDevice), "mc pos");
        }
        batch.evaluate(1);
        for (int k = 0; k < ns; ++k) {
            const float e_new = batch.potential(k);
            const float lb = dl[k] - (1.f / batch.sys[k].temperature) * (e_new - e_old[k]);
            bool accept = lb >= 0.f;
            if (!accept) {
                uint32_t key[4] = {batch.desc[k].seed, 2u, 0u, 0u};
                uint32_t ctr[4] = {(uint32_t)(round & 0xffffffffull), (uint32_t)(round >> 32), 0u, 1u};
                uint32_t bits[4];
                threefry_host(ctr, key, bits);
                accept = std::exp(lb) >= u01_open_closed(bits[0]);
            }
            mc_total[k][1]++;
            mc_frame[k][1]++;
            if (accept) {
                mc_total[k][0]++;
                mc_frame[k][0]++;
            } else {
                const Topology& T = batch.sys[k].prog->T;
                hip_check(hipMemcpy(batch.sys[k].arena + T.nodes[0].out, keep[k].data(), keep[k].size() * 4,
                                    hipMemcpyHostToDevice), "mc restore");
            }
        }
    }

    // One frame of every system at the current state: the potential and
    // derivative are evaluated first, as src/main.cpp:946-949 does before
    // collect_samples.
    void log_frame() {
        if (recenter)
            for (auto& s : batch.sys) {
                const Topology& T = s.prog->T;
                hip_check(upside_launch_recenter(s.arena + T.nodes[0].out, T.n_atom, recenter == 2, batch.stream),
                          "recenter");
            }
        batch.evaluate(1);
        for (size_t k = 0; k < batch.sys.size(); ++k) {
            const Topology& T = batch.sys[k].prog->T;
            vector<float> pos(T.n_atom * 3), mom(T.n_atom * 3);
            hip_check(hipMemcpy(pos.data(), batch.sys[k].arena + T.nodes[0].out, pos.size() * 4,
                                hipMemcpyDeviceToHost), "pos");
            hip_check(hipMemcpy(mom.data(), batch.sys[k].arena + T.mom, mom.size() * 4, hipMemcpyDeviceToHost),
                      "mom");
            double kin = 0.0;  // (1/2) <|p|^2> per atom, summed in double (src/main.cpp:822-826)
            for (int a = 0; a < T.n_atom; ++a) {
                float x = mom[a * 3], y = mom[a * 3 + 1], z = mom[a * 3 + 2];
                kin += x * x + y * y + z * z;
            }
            kin *= 0.5 / T.n_atom;
            const float t = 3 * batch.dt * (float)batch.round_num;  // inner_step*dt*round_num (:830-831)
            OutputLog& o = *out[k];
            if (o.hbond >= 0) {  // hbond value of every donor and acceptor site (column 6)
                const Node& hn = T.nodes[o.hb_node];
                vector<float> rows((size_t)hn.n_elem * hn.width);
                hip_check(hipMemcpy(rows.data(), batch.sys[k].arena + hn.out, rows.size() * 4, hipMemcpyDeviceToHost),
                          "hbond");
                o.hb_values.resize(hn.n_elem);
                for (int v = 0; v < hn.n_elem; ++v) o.hb_values[v] = rows[(size_t)v * hn.width + 6];
            }
            if (o.bad_solves >= 0) {
                int st4[4];
                hip_check(hipMemcpy(st4, batch.sys[k].iarena + o.rot_stats, sizeof(st4), hipMemcpyDeviceToHost),
                          "stats");
                o.n_bad = st4[1];
            }
            if (!o.rot_sets.empty()) {  // BP state of the evaluation above
                const char* split[3] = {"rotamer_free_energy", "rotamer_potential", "rotamer_entropy"};
                RotFetch fetched;  // one download of the BP state for the four values
                for (int q = 0; q < 3; ++q) o.rot_rows[q] = rotamer_values(batch, (int)k, o.rot_op, split[q], &fetched);
                vector<float> one = rotamer_values(batch, (int)k, o.rot_op, "rotamer_1body_energy", &fetched);
                for (int q = 0; q < o.rot_nprob; ++q) {
                    o.rot_rows[3 + q].resize(o.rot_nn);
                    for (int x = 0; x < o.rot_nn; ++x) o.rot_rows[3 + q][x] = one[(size_t)x * o.rot_nprob + q];
                }
            }
            for (auto& nl : o.node_logs) {
                const Node& nd = T.nodes[nl.node];
                const OpRec& op = T.ops[nd.op];
                nl.row.assign(nl.n, 0.f);
                float* A = batch.sys[k].arena;
                if (nl.kind == OutputLog::NL_RAMA) {  // per-residue potentials the kernel left after its values
                    hip_check(hipMemcpy(nl.row.data(), A + op.scratch + 2 * nl.n, nl.n * 4, hipMemcpyDeviceToHost),
                              "rama log");
                } else if (nl.kind == OutputLog::NL_SIGMOID) {
                    // scale * compact_sigmoid(sum_a w[t][a] coverage[r][a] - center[t], sharpness[t])
                    const int nr = op.ia[0], nt = op.ia[6];
                    vector<float> x((size_t)nr * nt);
                    hip_check(hipMemcpy(x.data(), A + op.in_out[0], x.size() * 4, hipMemcpyDeviceToHost), "coupling log");
                    const int* ct = T.ci.data() + op.ia[1];
                    const float *W = T.cf.data() + op.ia[2], *sc = T.cf.data() + op.ia[3];
                    const float *ce = T.cf.data() + op.ia[4], *sh = T.cf.data() + op.ia[5];
                    for (int r = 0; r < nr; ++r) {
                        const int t = ct[r];
                        float wn = 0.f;
                        for (int a = 0; a < nt; ++a) wn += W[t * nt + a] * x[(size_t)r * nt + a];
                        const float y = (wn - ce[t]) * sh[t];
                        float v = 0.25f * (y + 2.f) * (y - 1.f) * (y - 1.f);
                        if (y < -1.f) v = 1.f;
                        if (y > 1.f) v = 0.f;
                        nl.row[r] = v * sc[t];
                    }
                } else if (nl.kind == OutputLog::NL_OUTPUT) {
                    hip_check(hipMemcpy(nl.row.data(), A + nd.out, nl.n * 4, hipMemcpyDeviceToHost), "node log");
                } else {
                    vector<float> rows((size_t)nd.n_elem * nd.width);
                    hip_check(hipMemcpy(rows.data(), A + nd.out, rows.size() * 4, hipMemcpyDeviceToHost), "virtual log");
                    for (int v = 0; v < nd.n_elem; ++v)
                        for (int d = 0; d < 3; ++d) nl.row[(size_t)v * 3 + d] = rows[(size_t)v * nd.width + d];
                }
            }
            if (o.replica_index >= 0) {
                o.rep_index = replica_of[k];
                o.rep_swaps.clear();
                for (auto& sp : rex_partner[k]) {
                    o.rep_swaps.push_back((int)rex_stats[sp[0]][sp[1]][0]);
                    o.rep_swaps.push_back((int)rex_stats[sp[0]][sp[1]][1]);
                }
            }
            int st[2] = {0, 0};
            if (k < mc_frame.size()) { st[0] = mc_frame[k][0]; st[1] = mc_frame[k][1]; mc_frame[k] = {0, 0}; }
            out[k]->frame(pos.data(), mom.data(), T.n_atom, kin, (double)batch.potential((int)k), (double)t,
                          (double)batch.sys[k].temperature, st);
        }
    }
};

#define ABI_TRY try {
#define ABI_CATCH(ret)                                        \
    }                                                         \
    catch (const std::exception& e) {                         \
        fprintf(stderr, "\n\nERROR: %s\n", e.what());         \
        return ret;                                           \
    }                                                         \
    catch (...) {                                             \
        fprintf(stderr, "\n\nERROR: unknown error\n");        \
        return ret;                                           \
    }

extern "C" {

DerivEngine* construct_deriv_engine(int n_atom, const char* potential_file, bool quiet) {
    ABI_TRY
    (void)quiet;
    std::unique_ptr<DerivEngine> e(new DerivEngine);
    e->batch.create({string(potential_file)}, {1.f}, n_atom);
    if (e->batch.sys[0].prog->T.n_atom != n_atom) fail("n_atom mismatch");
    return e.release();
    ABI_CATCH(nullptr)
}

void free_deriv_engine(DerivEngine* engine) { delete engine; }

int evaluate_energy(float* energy, DerivEngine* engine, const float* pos) {
    ABI_TRY
    engine->batch.copy_pos_in(0, pos);
    engine->batch.evaluate(1);
    engine->evaluated = true;
    *energy = engine->batch.potential(0);
    return 0;
    ABI_CATCH(1)
}

int evaluate_deriv(float* deriv, DerivEngine* engine, const float* pos) {
    ABI_TRY
    auto& B = engine->batch;
    B.copy_pos_in(0, pos);
    B.evaluate(1);
    engine->evaluated = true;
    const Topology& T = B.sys[0].prog->T;
    hip_check(hipMemcpy(deriv, B.sys[0].arena + T.nodes[0].sens, T.n_atom * 3 * sizeof(float), hipMemcpyDeviceToHost), "deriv");
    return 0;
    ABI_CATCH(1)
}

// src/engine_c_library.cpp:68-78: replace a node's parameters.  Pair graphs
// (InteractionGraph::set_param, src/interaction_graph.h:417-426) where the
// pair-list cutoff does not depend on the table (the spline graphs and
// protein_hbond, whose cutoffs are fixed, src/bead_interaction.h:193-195,
// src/hbond.cpp:164-166, :258-260); the environment graphs, whose cutoff the
// reference recomputes from a new table, return nonzero.  The sigmoid
// couplings and hbond energy (src/environment.cpp:941-955, :1110-1115,
// src/hbond.cpp:722-737) take the layout get_param returns.
int set_param(int n_param, const float* param, DerivEngine* engine, const char* node_name) {
    ABI_TRY
    auto& B = engine->batch;
    Program& P = *B.sys[0].prog;
    const Topology& T = P.T;
    const Node* n = B.node(0, node_name);
    const int node = (int)(n - T.nodes.data());
    const vector<float> cur = P.get_params(node);
    if (cur.empty()) fail(string("set_param: no parameters exposed for node ") + node_name);
    vector<float> v(param, param + std::max(n_param, 0));
    for (size_t g = 0; g < T.graphs.size(); ++g) {
        if (T.graph_node[g] != node) continue;
        const PairGraph& G = T.graphs[g];
        if (G.kind != EDGE_ROTAMER && G.kind != EDGE_HBOND_COVERAGE && G.kind != EDGE_PROTEIN_HBOND)
            fail(string("set_param: the pair-list cutoff of ") + node_name + " depends on its parameters; not supported");
        if (n_param != (int)cur.size())
            fail("Bad param size, got " + std::to_string(n_param) + " params, but expected " +
                 std::to_string(cur.size()) + " params of shape (" + std::to_string(G.n_type1) + ", " +
                 std::to_string(G.n_type2) + ", " + std::to_string(G.n_param) + ")");
        if (G.sym) {
            // a symmetric table must equal its transpose with the angular
            // halves swapped (PosQuadSplineInteraction::is_compatible,
            // src/bead_interaction.h:211-220, checked by update_cutoffs,
            // :385-396); checked here before anything is replaced
            const int nka = 15, np = G.n_param, nt = G.n_type2;
            for (int a = 0; a < G.n_type1; ++a)
                for (int b = 0; b < nt; ++b) {
                    const float* p1 = v.data() + (a * nt + b) * np;
                    const float* p2 = v.data() + (b * nt + a) * np;
                    for (int k = 0; k < nka; ++k)
                        if (p1[k] != p2[k + nka] || p1[k + nka] != p2[k]) fail("bad angular match");
                    for (int k = 2 * nka; k < np; ++k)
                        if (p1[k] != p2[k]) fail("incompatible parameters");
                }
        }
    }
    if (n_param != (int)cur.size())
        fail(string("set_param: ") + node_name + " has " + std::to_string(cur.size()) + " parameters, got " +
             std::to_string(n_param));
    hip_check(hipStreamSynchronize(B.stream), "sync");
    P.put_params(node, v);
    P.param_set[node] = v;
    hip_check(hipMemcpy(P.d_cf, P.T.cf.data(), P.T.cf.size() * sizeof(float), hipMemcpyHostToDevice), "set_param");
    if (n->op >= 0)
        for (size_t k = 0; k < P.prog.size(); ++k)
            if (P.prog_node[k] == node) {
                std::copy(P.T.ops[n->op].fa, P.T.ops[n->op].fa + 16, P.prog[k].fa);
                hip_check(hipMemcpy(P.d_ops + k, &P.prog[k], sizeof(OpRec), hipMemcpyHostToDevice), "set_param");
            }
    engine->evaluated = false;
    return 0;
    ABI_CATCH(1)
}

// src/engine_c_library.cpp:94-110 and InteractionGraph::get_param_deriv
// (src/interaction_graph.h:406-416): d(potential)/d(interaction table) at the
// last evaluation, as sum over present edges of the edge sensitivity times the
// edge value's coefficient derivative, each edge adding to its type pair's row
// (and, in a symmetric graph, to the transposed row too, :544-547).  Provided
// for the spline pair graphs the reference builds it for in ff_2.1: rotamer
// pairs (src/bead_interaction.h:206-209) and hbond coverage
// (src/hbond.cpp:286-291).  Other nodes return nonzero.
int get_param_deriv(int n_param, float* deriv, DerivEngine* engine, const char* node_name) {
    ABI_TRY
    auto& B = engine->batch;
    if (!engine->evaluated) fail("get_param_deriv: evaluate the engine first");
    const Topology& T = B.sys[0].prog->T;
    const Node* n = B.node(0, node_name);
    const int op_kind = n->op >= 0 ? T.ops[n->op].kind : -1;
    if (op_kind == OP_SIGMOID_COUPLING || op_kind == OP_BB_SIGMOID_COUPLING || op_kind == OP_HBOND_ENERGY) {
        const OpRec& op = T.ops[n->op];
        const System& s = B.sys[0];
        const Program& P = *s.prog;
        CouplingDerivArgs a;
        std::memset(&a, 0, sizeof(a));
        a.kind = op.kind;
        const Node& p0 = T.nodes[n->parents[0]];
        a.cov = s.arena + p0.out;
        if (op.kind == OP_HBOND_ENERGY) {
            const Node& p1 = T.nodes[n->parents[1]];
            a.nt = 1;
            a.n_res = op.ia[0];
            a.pp = s.arena + p0.out;
            a.wp = p0.width;
            a.rama = s.arena + p1.out;
            a.wr = p1.width;
            a.h1o = P.d_ci + op.ia[1];
            a.h1i = P.d_ci + op.ia[2];
            a.h2o = P.d_ci + op.ia[3];
            a.h2i = P.d_ci + op.ia[4];
            for (int k = 0; k < 12; ++k) a.hp[k] = op.fa[k];
            a.n_out = 12;
        } else if (op.kind == OP_SIGMOID_COUPLING) {
            a.nt = op.ia[6];
            a.n_res = op.ia[0];
            a.n_indep = op.ia[7];
            if (a.n_indep != 1 && a.n_indep != 20 && a.n_indep != 400)
                fail("the number of independent weights should be 1, 20 or 400");
            a.ctype = P.d_ci + op.ia[1];
            a.weights = P.d_cf + op.ia[2];
            a.scale = P.d_cf + op.ia[3];
            a.center = P.d_cf + op.ia[4];
            a.sharp = P.d_cf + op.ia[5];
            a.n_out = 3 * a.nt + a.nt * a.nt;
        } else {
            a.nt = op.ia[0];
            a.n_res = op.ia[2];
            a.weights = P.d_cf + op.ia[1];
            a.cov_hb = s.arena + T.nodes[n->parents[1]].out;
            a.bb_center = op.fa[1];
            a.bb_sharp = op.fa[2];
            a.bb_hbond_weight = op.fa[3];
            a.n_out = 4;
        }
        if (n_param != a.n_out)
            fail("Wrong number of parameters, expected " + std::to_string(a.n_out) + " but got " + std::to_string(n_param));
        float* d_out = nullptr;
        hip_check(hipMalloc((void**)&d_out, (size_t)a.n_out * sizeof(float)), "param deriv");
        a.out = d_out;
        hipError_t e = upside_launch_coupling_param_deriv(a, B.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(B.stream);
        if (e == hipSuccess) e = hipMemcpy(deriv, d_out, (size_t)a.n_out * sizeof(float), hipMemcpyDeviceToHost);
        (void)hipFree(d_out);
        hip_check(e, "param deriv");
        return 0;
    }
    for (size_t g = 0; g < T.graphs.size(); ++g) {
        if (&T.nodes[T.graph_node[g]] != n) continue;
        const PairGraph& G = T.graphs[g];
        // protein_hbond and backbone_pairs have no hook in the reference (its
        // C ABI then reports a size mismatch): an error here too
        if (G.kind != EDGE_ROTAMER && G.kind != EDGE_HBOND_COVERAGE && G.kind != EDGE_ENV_WEIGHTED &&
            G.kind != EDGE_ENV_HBOND)
            fail(string("get_param_deriv: no parameter derivative for node ") + node_name);
        const int rows = G.n_type1 * G.n_type2;
        const int size = rows * G.n_param;
        if (size != n_param)
            fail("Wrong number of parameters, expected " + std::to_string(size) + " but got " + std::to_string(n_param));
        if (G.kind == EDGE_ENV_WEIGHTED || G.kind == EDGE_ENV_HBOND) {
            // the reference's environment-coverage interactions have no
            // parameter derivative: zeros (src/environment.cpp:164-167, :222-225)
            std::fill(deriv, deriv + size, 0.f);
            return 0;
        }
        const int spline_params = G.kind == EDGE_ROTAMER ? 2 * 15 + 2 * 12 : 2 * 15 + 2 * 10;
        if (G.n_param != spline_params)
            fail(string("get_param_deriv: ") + node_name + " has " + std::to_string(G.n_param) +
                 " parameters per type pair, the spline needs " + std::to_string(spline_params));
        const System& s = B.sys[0];
        const Program& P = *s.prog;
        ParamDerivArgs a;
        a.x1 = s.arena + G.out1;
        a.x2 = s.arena + G.out2;
        a.w1 = G.w1;
        a.w2 = G.w2;
        a.loc1 = P.d_ci + G.loc1;
        a.loc2 = P.d_ci + G.loc2;
        a.type1 = P.d_ci + G.type1;
        a.type2 = P.d_ci + G.type2;
        a.row = s.iarena + G.row;
        a.nbr = s.iarena + G.nbr;
        a.act = s.iarena + G.act;
        a.es = s.arena + G.esens;
        int st[4];
        hip_check(hipMemcpy(st, s.iarena + G.state, sizeof(st), hipMemcpyDeviceToHost), "state");
        a.n_slot = st[3];
        a.param = P.d_cf + G.param;
        a.n_type2 = G.n_type2;
        a.n_param = G.n_param;
        a.sym = G.sym;
        a.kind = G.kind;
        float* d_out = nullptr;
        hip_check(hipMalloc((void**)&d_out, (size_t)size * sizeof(float)), "param deriv");
        a.out = d_out;
        hipError_t e = upside_launch_param_deriv(a, rows, B.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(B.stream);
        if (e == hipSuccess) e = hipMemcpy(deriv, d_out, (size_t)size * sizeof(float), hipMemcpyDeviceToHost);
        (void)hipFree(d_out);
        hip_check(e, "param deriv");
        return 0;
    }
    fail(string("get_param_deriv: no parameters exposed for node ") + node_name);
    ABI_CATCH(1)
}

int get_param(int n_param, float* param, DerivEngine* engine, const char* node_name) {
    ABI_TRY
    auto& B = engine->batch;
    const Program& P = *B.sys[0].prog;
    const Node* n = B.node(0, node_name);
    const vector<float> v = P.get_params((int)(n - P.T.nodes.data()));
    if (v.empty()) fail(string("no parameters exposed for node ") + node_name);
    if ((int)v.size() != n_param)
        fail("Wrong number of parameters, expected " + std::to_string(v.size()) + " but got " + std::to_string(n_param));
    std::copy(v.begin(), v.end(), param);
    return 0;
    ABI_CATCH(1)
}

int get_output_dims(int* n_elem, int* elem_width, DerivEngine* engine, const char* node_name) {
    ABI_TRY
    const Node* n = engine->batch.node(0, node_name);
    if (n->potential) { *n_elem = 1; *elem_width = 1; }
    else { *n_elem = n->n_elem; *elem_width = n->width; }
    return 0;
    ABI_CATCH(1)
}

static int copy_node(int n_output, float* output, DerivEngine* engine, const char* node_name, bool sens) {
    ABI_TRY
    auto& B = engine->batch;
    const Node* n = B.node(0, node_name);
    const Topology& T = B.sys[0].prog->T;
    if (n->potential) {
        if (n_output != 1) fail("wrong size for potential node");
        hip_check(hipMemcpy(output, B.sys[0].arena + T.pot + n->pot_slot, 4, hipMemcpyDeviceToHost), "potential");
        return 0;
    }
    if (n_output != n->n_elem * n->width) fail("wrong size for CoordNode");
    hip_check(hipMemcpy(output, B.sys[0].arena + (sens ? n->sens : n->out), (size_t)n_output * 4, hipMemcpyDeviceToHost), "node copy");
    return 0;
    ABI_CATCH(1)
}

int get_output(int n_output, float* output, DerivEngine* engine, const char* node_name) {
    return copy_node(n_output, output, engine, node_name, false);
}
int get_sens(int n_output, float* output, DerivEngine* engine, const char* node_name) {
    return copy_node(n_output, output, engine, node_name, true);
}

// Rotamer BP state read back after an evaluation, in the reference's layouts
// (src/rotamer.cpp:736-834): per-residue values in bead order (arrange_energies,
// :1023-1055), graph-ordered values indexed by node class (1, 3, then 6
// rotamers) and the residue's index within its class, which is the bead id's
// bits above the rotamer fields (py/upside_config.py:1171-1178).
static vector<float> rotamer_values(Batch& B, int sys, int op, const string& ln, RotFetch* cache) {
    const Topology& T = B.sys[sys].prog->T;
    const RotamerDev& R = T.rots[T.ops[op].rot];
    const PairGraph& G = T.graphs[T.ops[op].graph];
    const float* A = B.sys[sys].arena;
    const int* I = B.sys[sys].iarena;
    const int nn = R.n_node, cap = R.ccap;
    const int* nrot = T.ci.data() + R.node_nrot;
    const int* bead0 = T.ci.data() + R.node_bead0;
    const int* brot = T.ci.data() + R.bead_rot;
    const int* bnode = T.ci.data() + R.bead_node;
    const int* ids = T.ci.data() + G.id1;
    auto fetch = [](const float* src, size_t n) {
        vector<float> v(n);
        if (n) hip_check(hipMemcpy(v.data(), src, n * 4, hipMemcpyDeviceToHost), "rotamer state");
        return v;
    };
    auto ifetch = [](const int* src, size_t n) {
        vector<int> v(n);
        if (n) hip_check(hipMemcpy(v.data(), src, n * 4, hipMemcpyDeviceToHost), "rotamer state");
        return v;
    };
    RotFetch local;
    RotFetch& F = cache ? *cache : local;
    if (!F.filled) {
        F.prob = fetch(A + R.nprob, (size_t)nn * 6);
        F.cur = fetch(A + R.ncur, (size_t)nn * 6);
        F.off = fetch(A + R.noff, nn);
        int st[8];
        hip_check(hipMemcpy(st, I + R.stats, sizeof(st), hipMemcpyDeviceToHost), "stats");
        F.nbp = st[4];
        F.bpf = ifetch(I + R.bpf, F.nbp);
        F.bps = ifetch(I + R.bps, F.nbp);
        // matrices: 36 contiguous entries r1*6 + r2 per BP edge; marginals: rows
        // q = r1*6 + r2 of stride cap over the edges (one copy of all rows)
        F.P = fetch(A + R.P, (size_t)36 * F.nbp);
        F.M.assign((size_t)36 * F.nbp, 0.f);
        const vector<float> rows = fetch(A + R.marg, (size_t)36 * cap);
        for (int q = 0; q < 36; ++q)
            for (int e = 0; e < F.nbp; ++e) F.M[(size_t)e * 36 + q] = rows[(size_t)q * cap + e];
        // 1-rotamer residue pairs: one bead pair each, probability exp(-E)
        int gst[4];
        hip_check(hipMemcpy(gst, I + G.state, sizeof(gst), hipMemcpyDeviceToHost), "state");
        const int E = gst[3];
        vector<int> row = ifetch(I + G.row, E), nbr = ifetch(I + G.nbr, E), act = ifetch(I + G.act, E);
        vector<float> val = fetch(A + G.val, E);
        for (int s = 0; s < E; ++s) {
            if (!act[s]) continue;
            int a = bnode[row[s]], b = bnode[nbr[s]];
            if (nrot[a] != 1 || nrot[b] != 1) continue;
            F.e11.push_back({a, b});
            F.p11.push_back(expf(-val[s]));
        }
        F.filled = true;
    }
    const vector<float>& prob = F.prob;
    const vector<float>& cur = F.cur;
    const vector<float>& off = F.off;
    const vector<float>& P = F.P;
    const vector<float>& M = F.M;
    const vector<float>& p11 = F.p11;
    const vector<int>& bpf = F.bpf;
    const vector<int>& bps = F.bps;
    const vector<std::array<int, 2>>& e11 = F.e11;
    const int nbp = F.nbp;

    // per-residue split of the Bethe free energy: residue_free_energies
    // (:929-963) and its average-energy and entropy parts residue_entropy<0>
    // and <1> (:965-1003, node_entropy :303-321, edge_entropy :473-497)
    const int split = ln == "rotamer_free_energy" ? 2 : ln == "rotamer_potential" ? 0 : ln == "rotamer_entropy" ? 1 : -1;
    if (split >= 0) {
        vector<float> en(nn, 0.f);
        for (int x = 0; x < nn; ++x) {
            float s = 0.f, e = split == 1 ? 0.f : off[x];
            for (int r = 0; r < nrot[x]; ++r) s += cur[x * 6 + r];
            for (int r = 0; r < nrot[x]; ++r) {
                float b = cur[x * 6 + r] / s, pr = prob[x * 6 + r];
                if (split == 2) e += b * logf((1e-10f + b) / (1e-10f + pr));
                else if (split == 0) e -= b * logf(1e-10f + pr);
                else e += b * logf(1e-10f + b);
            }
            en[x] += e;
        }
        if (split != 1)
            for (size_t k = 0; k < e11.size(); ++k) {
                float e = -logf(p11[k]);
                en[e11[k][0]] += 0.5f * e;
                en[e11[k][1]] += 0.5f * e;
            }
        for (int e = 0; e < nbp; ++e) {
            int f = bpf[e], s2 = bps[e];
            float fe = 0.f;
            for (int i = 0; i < nrot[f]; ++i)
                for (int j = 0; j < nrot[s2]; ++j) {
                    float p = M[(size_t)e * 36 + i * 6 + j], pr = P[(size_t)e * 36 + i * 6 + j];
                    float bb = cur[f * 6 + i] * cur[s2 * 6 + j];
                    if (split == 2) fe += p * logf((1e-10f + p) / (1e-10f + pr * bb));
                    else if (split == 0) fe -= p * logf(1e-10f + pr);
                    else fe += p * logf((1e-10f + p) / (1e-10f + bb));
                }
            en[f] += 0.5f * fe;
            en[s2] += 0.5f * fe;
        }
        return en;
    }
    if (ln == "rotamer_1body_energy") {  // rotamer_1body_energy (:1005-1021), [residue][energy node]
        vector<float> res((size_t)nn * R.n_prob, 0.f);
        const vector<int> loc1(T.ci.data() + G.loc1, T.ci.data() + G.loc1 + G.n1);
        for (int q = 0; q < R.n_prob; ++q) {
            int maxloc = 0;
            for (int b = 0; b < G.n1; ++b) maxloc = std::max(maxloc, loc1[b]);
            vector<float> e1 = fetch(A + R.prob_out[q], (size_t)(maxloc + 1) * R.prob_w[q]);
            for (int b = 0; b < G.n1; ++b) {
                int x = bnode[b];
                res[(size_t)x * R.n_prob + q] += cur[x * 6 + brot[b]] * e1[(size_t)loc1[b] * R.prob_w[q]];
            }
        }
        return res;
    }
    // graph order: index of each node among all nodes as the reference numbers them
    int n_class[7] = {0};
    for (int x = 0; x < nn; ++x) n_class[nrot[x]] += 1;
    const int start[7] = {0, 0, 0, n_class[1], 0, 0, n_class[1] + n_class[3]};
    vector<int> gi(nn);
    for (int x = 0; x < nn; ++x) {
        int local = ids[bead0[x]] >> 8;
        if (local < 0 || local >= n_class[nrot[x]])
            fail("rotamer: residue index within its rotamer class out of range for graph-order values");
        gi[x] = start[nrot[x]] + local;
    }
    if (ln == "node_energy") {  // -log of the node probability, 1e5 past the rotamer count
        vector<float> res((size_t)nn * 6);
        for (int x = 0; x < nn; ++x)
            for (int r = 0; r < 6; ++r) res[(size_t)gi[x] * 6 + r] = r < nrot[x] ? -logf(prob[x * 6 + r]) : 1e5f;
        return res;
    }
    // edge_energy / edge_marginal_in_graph_order: [n_node][n_node][6][6]
    const bool marginal = ln == "edge_marginal_in_graph_order";
    vector<float> res((size_t)nn * nn * 36, 0.f);
    auto at = [&](int i1, int i2, int r1, int r2) -> float& {
        return res[(((size_t)i1 * nn + i2) * 6 + r1) * 6 + r2];
    };
    if (marginal) {  // residue pairs without a BP edge are uncorrelated
        vector<float> nm((size_t)nn * 6, 0.f);
        for (int x = 0; x < nn; ++x)
            for (int r = 0; r < nrot[x]; ++r) nm[(size_t)gi[x] * 6 + r] = cur[x * 6 + r];
        for (int i1 = 0; i1 < nn; ++i1)
            for (int i2 = 0; i2 < nn; ++i2)
                for (int r1 = 0; r1 < 6; ++r1)
                    for (int r2 = 0; r2 < 6; ++r2)
                        at(i1, i2, r1, r2) = i1 == i2 ? nm[i1 * 6 + r1] * (r1 == r2) : nm[i1 * 6 + r1] * nm[i2 * 6 + r2];
    }
    for (size_t k = 0; k < e11.size(); ++k) {  // the single state pair: marginal 1
        int i1 = gi[e11[k][0]], i2 = gi[e11[k][1]];
        at(i1, i2, 0, 0) = at(i2, i1, 0, 0) = marginal ? 1.f : -logf(p11[k]);
    }
    for (int e = 0; e < nbp; ++e) {
        int f = bpf[e], s2 = bps[e], i1 = gi[f], i2 = gi[s2];
        for (int r1 = 0; r1 < nrot[f]; ++r1)
            for (int r2 = 0; r2 < nrot[s2]; ++r2) {
                float v = marginal ? M[(size_t)e * 36 + r1 * 6 + r2] : -logf(P[(size_t)e * 36 + r1 * 6 + r2]);
                at(i1, i2, r1, r2) = at(i2, i1, r2, r1) = v;
            }
    }
    return res;
}

int get_value_by_name(int n_output, float* output, DerivEngine* engine, const char* node_name, const char* log_name) {
    ABI_TRY
    auto& B = engine->batch;
    const Topology& T = B.sys[0].prog->T;
    const Node* n = B.node(0, node_name);
    int gidx = -1;
    for (size_t g = 0; g < T.graphs.size(); ++g)
        if (&T.nodes[T.graph_node[g]] == n) gidx = (int)g;
    string ln(log_name);
    vector<float> out;
    if (gidx >= 0 && ln == "graph_shape") {
        const PairGraph& G = T.graphs[gidx];
        out = {(float)G.n1, (float)G.dim1, (float)G.n2, (float)G.dim2};
    } else if (n->type == "rotamer" && ln == "bp_edge_classes") {
        // present residue-pair edges by rotamer-count class: 11 13 16 33 36 66
        const RotamerDev& R = T.rots[T.ops[n->op].rot];
        const PairGraph& G = T.graphs[T.ops[n->op].graph];
        int* I = B.sys[0].iarena;
        int st[4];
        hip_check(hipMemcpy(st, I + G.state, sizeof(st), hipMemcpyDeviceToHost), "state");
        int E = st[3];
        vector<int> act(E), scand(E), row(E), nbr(E);
        if (E) {
            hip_check(hipMemcpy(act.data(), I + G.act, E * 4, hipMemcpyDeviceToHost), "bp");
            hip_check(hipMemcpy(scand.data(), I + R.slot_cand, E * 4, hipMemcpyDeviceToHost), "bp");
            hip_check(hipMemcpy(row.data(), I + G.row, E * 4, hipMemcpyDeviceToHost), "bp");
            hip_check(hipMemcpy(nbr.data(), I + G.nbr, E * 4, hipMemcpyDeviceToHost), "bp");
        }
        const int* nrot = T.ci.data() + R.node_nrot;
        const int* bnode = T.ci.data() + R.bead_node;
        out.assign(6, 0.f);
        auto cls = [](int a, int b) {
            if (a > b) std::swap(a, b);
            if (a == 1) return b == 1 ? 0 : (b == 3 ? 1 : 2);
            return a == 3 ? (b == 3 ? 3 : 4) : 5;
        };
        std::set<int> seen;
        for (int s2 = 0; s2 < E; ++s2)
            if (act[s2] && seen.insert(scand[s2]).second)
                out[cls(nrot[bnode[row[s2]]], nrot[bnode[nbr[s2]]])] += 1.f;
    } else if (n->type == "rotamer" && ln == "node_classes") {
        const RotamerDev& R = T.rots[T.ops[n->op].rot];
        const int* nrot = T.ci.data() + R.node_nrot;
        out.assign(3, 0.f);
        for (int a = 0; a < R.n_node; ++a) out[nrot[a] == 1 ? 0 : (nrot[a] == 3 ? 1 : 2)] += 1.f;
    } else if (gidx >= 0 && (ln == "count_edges_by_type" || ln == "n_edge" || ln == "edge_pairs" || ln == "edge_values")) {
        const PairGraph& G = T.graphs[gidx];
        float* A = B.sys[0].arena;
        int* I = B.sys[0].iarena;
        int st[4];
        hip_check(hipMemcpy(st, I + G.state, sizeof(st), hipMemcpyDeviceToHost), "state");
        int E = st[3];
        vector<int> row(E), nbr(E), act(E);
        vector<float> val(E);
        if (E) {
            hip_check(hipMemcpy(row.data(), I + G.row, E * 4, hipMemcpyDeviceToHost), "edges");
            hip_check(hipMemcpy(nbr.data(), I + G.nbr, E * 4, hipMemcpyDeviceToHost), "edges");
            hip_check(hipMemcpy(act.data(), I + G.act, E * 4, hipMemcpyDeviceToHost), "edges");
            hip_check(hipMemcpy(val.data(), A + G.val, E * 4, hipMemcpyDeviceToHost), "edges");
        }
        const int* t1 = T.ci.data() + G.type1;
        const int* t2 = T.ci.data() + G.type2;
        int n_type1 = 0;
        for (int i = 0; i < G.n1; ++i) n_type1 = std::max(n_type1, t1[i] + 1);
        vector<float> by_type((size_t)n_type1 * G.n_type2, 0.f);
        vector<float> pairs, vals;
        for (int s2 = 0; s2 < E; ++s2) {  // slots are in (row, partner) order
            if (!act[s2]) continue;
            int i = row[s2], j = nbr[s2];
            by_type[(size_t)t1[i] * G.n_type2 + t2[j]] += 1.f;
            pairs.push_back((float)i);
            pairs.push_back((float)j);
            vals.push_back(val[s2]);
        }
        if (ln == "count_edges_by_type") out = by_type;
        else if (ln == "n_edge") out = {(float)vals.size()};
        else if (ln == "edge_pairs") out = pairs;
        else out = vals;
    } else if (n->type == "rotamer") {
        const RotamerDev& R = T.rots[T.ops[n->op].rot];
        int st[4];
        hip_check(hipMemcpy(st, B.sys[0].iarena + R.stats, sizeof(st), hipMemcpyDeviceToHost), "stats");
        if (ln == "n_node") out = {(float)R.n_node};
        else if (ln == "read n_bad_solve") out = {(float)st[1]};
        else if (ln == "read n_bad_solve and reset") {
            out = {(float)st[1]};
            st[1] = 0;
            hip_check(hipMemcpy(B.sys[0].iarena + R.stats, st, sizeof(st), hipMemcpyHostToDevice), "stats");
        } else if (ln == "iterations") out = {(float)st[0]};
        else if (ln == "rotamer_free_energy" || ln == "rotamer_1body_energy" || ln == "node_energy" ||
                 ln == "edge_energy" || ln == "edge_marginal_in_graph_order" || ln == "rotamer_potential" ||
                 ln == "rotamer_entropy")
            out = rotamer_values(B, 0, n->op, ln);
        else fail("Value " + ln + " not implemented");
    } else {
        fail("Value " + ln + " not implemented");
    }
    if (n_output != (int)out.size())
        fail("expected size (" + std::to_string(n_output) + " elements) inconsistent with actual size (" +
             std::to_string(out.size()) + ")");
    std::copy(out.begin(), out.end(), output);
    return 0;
    ABI_CATCH(1)
}

// ---- clamped spline helpers (src/engine_c_library.cpp:197-277, src/spline.cpp:158-190) ----
int clamped_spline_solve(int N, float* bspline_coeff, const float* values) {
    // Solve for N B-spline coefficients whose spline takes values[0..N-3] at
    // knots 1..N-2 with zero end derivatives (coeff[0]=coeff[2], coeff[N-1]=coeff[N-3]).
    if (N < 4) return 1;
    int n = N - 2;
    vector<double> a(n, 1. / 6.), b(n, 2. / 3.), c(n, 1. / 6.), d(values, values + n);
    a[n - 1] *= 2.;  // fold the clamping conditions into the tridiagonal system
    c[0] *= 2.;
    for (int k = 1; k < n; ++k) {
        double m = a[k] / b[k - 1];
        b[k] -= m * c[k - 1];
        d[k] -= m * d[k - 1];
    }
    d[n - 1] /= b[n - 1];
    for (int k = n - 2; k >= 0; --k) d[k] = (d[k] - c[k] * d[k + 1]) / b[k];
    for (int k = 0; k < n; ++k) bspline_coeff[k + 1] = (float)d[k];
    bspline_coeff[0] = bspline_coeff[2];
    bspline_coeff[N - 1] = bspline_coeff[N - 3];
    return 0;
}

static void clamped_eval(const float* p, float x, int n_knot, float& v, float& d) {
    if (x <= 1.f) { v = p[0] / 6.f + p[1] * (2.f / 3.f) + p[2] / 6.f; d = 0.f; return; }
    if (x >= n_knot - 2) { v = p[n_knot - 3] / 6.f + p[n_knot - 2] * (2.f / 3.f) + p[n_knot - 1] / 6.f; d = 0.f; return; }
    int xb = (int)x;
    float y = x - xb;
    const float* c = p + xb - 1;
    float omy = 1.f - y, y2 = y * y, y3 = y2 * y;
    v = (omy * omy * omy * c[0] + (3 * y3 - 6 * y2 + 4) * c[1] + (-3 * y3 + 3 * y2 + 3 * y + 1) * c[2] + y3 * c[3]) / 6.f;
    d = -0.5f * omy * omy * c[0] + (1.5f * y2 - 2 * y) * c[1] + (-1.5f * y2 + y + 0.5f) * c[2] + 0.5f * y2 * c[3];
}

int clamped_spline_value(int N, float* result, const float* bspline_coeff, int nx, float* x) {
    for (int i = 0; i < nx; ++i) { float d; clamped_eval(bspline_coeff, x[i], N, result[i], d); }
    return 0;
}

int get_clamped_value_and_deriv(int N, float* result, const float* bspline_coeff, int nx, float* x) {
    for (int i = 0; i < nx; ++i) clamped_eval(bspline_coeff, x[i], N, result[2 * i], result[2 * i + 1]);
    return 0;
}

int get_clamped_coeff_deriv(int N, float* result, const float* bspline_coeff, float x) {
    (void)bspline_coeff;
    for (int i = 0; i < N; ++i) result[i] = 0.f;
    if (x <= 1.f) { result[0] = 1.f / 6.f; result[1] = 2.f / 3.f; result[2] = 1.f / 6.f; return 0; }
    if (x >= N - 2) { result[N - 3] = 1.f / 6.f; result[N - 2] = 2.f / 3.f; result[N - 1] = 1.f / 6.f; return 0; }
    int xb = (int)x;
    float y = x - xb, omy = 1.f - y, y2 = y * y, y3 = y2 * y;
    result[xb - 1] = omy * omy * omy / 6.f;
    result[xb] = (3 * y3 - 6 * y2 + 4) / 6.f;
    result[xb + 1] = (-3 * y3 + 3 * y2 + 3 * y + 1) / 6.f;
    result[xb + 2] = y3 / 6.f;
    return 0;
}

// ---- batched extension ----
UpsideBatch* upside_batch_create(int n_sys, const char* const* up_paths, const float* temperatures,
                                 uint32_t base_seed, float dt, int thermostat_interval) {
    ABI_TRY
    std::unique_ptr<UpsideBatch> b(new UpsideBatch);
    vector<string> paths;
    vector<float> temps;
    for (int k = 0; k < n_sys; ++k) {
        paths.push_back(up_paths[k]);
        temps.push_back(temperatures ? temperatures[k] : 1.f);
    }
    b->batch.dt = dt;
    b->batch.thermostat_interval = std::max(1, thermostat_interval);
    b->batch.base_seed = base_seed;
    b->batch.create(paths, temps, 0);
    set_initial_positions(b->batch);
    b->batch.thermalize();
    return b.release();
    ABI_CATCH(nullptr)
}

void upside_batch_free(UpsideBatch* batch) { delete batch; }

int upside_batch_n_atom(UpsideBatch* b, int sys) {
    if (!b || sys < 0 || sys >= (int)b->batch.sys.size()) return -1;
    return b->batch.sys[sys].prog->T.n_atom;
}

static int batch_copy(UpsideBatch* b, int sys, float* host, bool to_device, bool mom) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("system index out of range");
    const Topology& T = B.sys[sys].prog->T;
    float* dev = B.sys[sys].arena + (mom ? T.mom : T.nodes[0].out);
    size_t bytes = (size_t)T.n_atom * 3 * sizeof(float);
    hip_check(hipMemcpy(to_device ? (void*)dev : (void*)host, to_device ? (const void*)host : (const void*)dev, bytes,
                        to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost), "copy");
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_pos(UpsideBatch* b, int sys, const float* pos) { return batch_copy(b, sys, (float*)pos, true, false); }
int upside_batch_get_pos(UpsideBatch* b, int sys, float* pos) { return batch_copy(b, sys, pos, false, false); }
int upside_batch_set_mom(UpsideBatch* b, int sys, const float* mom) { return batch_copy(b, sys, (float*)mom, true, true); }
int upside_batch_get_mom(UpsideBatch* b, int sys, float* mom) { return batch_copy(b, sys, mom, false, true); }

int upside_batch_open_output(UpsideBatch* b, const char* const* output_paths, int record_momentum, int log_level) {
    ABI_TRY
    auto& B = b->batch;
    const int ns = (int)B.sys.size();
    std::set<string> inputs;
    if (!output_paths)
        for (int k = 0; k < ns; ++k)
            if (!inputs.insert(B.sys[k].prog->path).second)
                fail("open_output: systems share the input file " + B.sys[k].prog->path + "; give output paths");
    for (auto& o : b->out) o->close();
    b->out.clear();
    b->record_mom = record_momentum != 0;
    for (int k = 0; k < ns; ++k) {
        b->out.emplace_back(new OutputLog());
        const string path = output_paths ? string(output_paths[k]) : B.sys[k].prog->path;
        vector<int> partners;  // the other system of each swap this one takes part in (:219-225)
        for (auto& sp : b->rex_partner.size() > (size_t)k ? b->rex_partner[k] : vector<std::array<int, 2>>()) {
            const auto& pr = b->rex_sets[sp[0]][sp[1]];
            partners.push_back(pr[0] != k ? pr[0] : pr[1]);
        }
        b->out.back()->open(path, !output_paths, B.sys[k].prog->T.n_atom, b->record_mom, b->mc_interval > 0,
                            B.sys[k].prog->T, log_level, b->rex_interval > 0 ? &partners : nullptr);
    }
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_monte_carlo(UpsideBatch* b, int mc_interval) {
    ABI_TRY
    if (mc_interval < 0) fail("set_monte_carlo: negative interval");
    if (mc_interval > 0)
        for (auto& s : b->batch.sys)
            if (s.prog->pivots.n_loc() == 0) fail("set_monte_carlo: " + s.prog->path + " has no /input/pivot_moves");
    b->mc_interval = mc_interval;
    b->mc_total.assign(b->batch.sys.size(), {0, 0});
    b->mc_frame.assign(b->batch.sys.size(), {0, 0});
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_recentering(UpsideBatch* b, int mode) {
    ABI_TRY
    if (mode < 0 || mode > 2) fail("set_recentering: mode must be 0 (none), 1 (x, y, z) or 2 (x, y)");
    b->recenter = mode;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_annealing(UpsideBatch* b, double factor, double start, double end, double duration) {
    ABI_TRY
    auto& B = b->batch;
    if (!(factor > 0.0)) fail("set_annealing: factor must be positive");
    // the reference's option handling (src/main.cpp:593-600); the run's total
    // duration, the CLI's default for the annealing duration, is the caller's
    if (start < 0.0) start = 0.0;
    if (end > start) duration = end - start;
    if (factor != 1.0 && !(duration > 0.0)) fail("set_annealing: give an end after the start or a positive duration");
    B.anneal = factor != 1.0;
    B.anneal_factor = factor;
    B.anneal_start = start;
    B.anneal_end = end;
    B.anneal_duration = duration;
    B.sync_temperature();
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_thermostat_timescale(UpsideBatch* b, float timescale) {
    ABI_TRY
    auto& B = b->batch;
    if (!(timescale > 0.f)) fail("set_thermostat_timescale: timescale must be positive");
    if (B.round_num != 0) fail("set_thermostat_timescale: only before the first round");
    B.thermostat_timescale = timescale;
    B.pull_desc();
    for (size_t k = 0; k < B.sys.size(); ++k) {
        const SysDesc d = B.make_desc((int)k);
        B.desc[k].mom_scale = d.mom_scale;
        B.desc[k].noise_scale = d.noise_scale;
        B.desc[k].n_invocations = 0;
    }
    B.push_desc();
    B.thermalize();  // the fresh-start thermalisation again, with the new parameters (src/main.cpp:800-803)
    return 0;
    ABI_CATCH(1)
}

int upside_batch_mc_stats(UpsideBatch* b, int sys, long long* accepted, long long* attempted) {
    ABI_TRY
    if (sys < 0 || sys >= (int)b->mc_total.size()) fail("mc_stats: no Monte Carlo state for this system");
    *accepted = b->mc_total[sys][0];
    *attempted = b->mc_total[sys][1];
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_replica_exchange(UpsideBatch* b, int replica_interval, int criterion, int n_sets,
                                      const int* set_sizes, const int* pairs) {
    ABI_TRY
    auto& B = b->batch;
    const int ns = (int)B.sys.size();
    if (replica_interval < 0) fail("set_replica_exchange: negative interval");
    if (criterion != 0 && criterion != 1) fail("replica exchange: exchange criterion must be 0 or 1");
    vector<vector<std::array<int, 2>>> sets;
    int at = 0;
    for (int s = 0; s < n_sets; ++s) {
        if (set_sizes[s] < 1) fail("set_replica_exchange: empty swap set");
        std::set<int> seen;
        sets.emplace_back();
        for (int p = 0; p < set_sizes[s]; ++p, ++at) {
            const int a = pairs[2 * at], c = pairs[2 * at + 1];
            if (a < 0 || a >= ns || c < 0 || c >= ns) fail("invalid system");
            if (a == c || seen.count(a) || seen.count(c))
                fail("Overlapping indices in swap set.  No replica index can appear more than once in a swap set.");
            seen.insert(a);
            seen.insert(c);
            sets.back().push_back({a, c});
        }
    }
    if (replica_interval > 0) {
        if (sets.empty()) fail("replica exchange requested but no swap sets proposed");
        for (int k = 1; k < ns; ++k)
            if (B.sys[k].prog->T.n_atom != B.sys[0].prog->T.n_atom)
                fail("Replica exchange requires all systems have the same number of atoms");
    }
    b->rex_interval = replica_interval;
    b->rex_criterion = criterion;
    b->rex_sets = sets;
    b->rex_stats.clear();
    for (auto& st : sets) b->rex_stats.emplace_back(st.size(), std::array<long long, 2>{0, 0});
    b->rex_partner.assign(ns, {});
    for (size_t s = 0; s < sets.size(); ++s)
        for (size_t p = 0; p < sets[s].size(); ++p)
            for (int x : {sets[s][p][0], sets[s][p][1]}) b->rex_partner[x].push_back({(int)s, (int)p});
    b->replica_of.resize(ns);
    for (int k = 0; k < ns; ++k) b->replica_of[k] = k;
    b->rex_attempts = 0;
    b->rex_next = replica_interval > 0 ? UpsideBatch::next_swap_round(B.round_num, replica_interval) : 0;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_replica_stats(UpsideBatch* b, int* replica_index, long long* successes, long long* attempts) {
    ABI_TRY
    for (size_t k = 0; k < b->replica_of.size(); ++k) replica_index[k] = b->replica_of[k];
    size_t i = 0;
    for (auto& st : b->rex_stats)
        for (auto& x : st) {
            successes[i] = x[0];
            attempts[i] = x[1];
            ++i;
        }
    return 0;
    ABI_CATCH(1)
}

int upside_batch_run_logged(UpsideBatch* b, int n_round, int frame_interval) {
    ABI_TRY
    auto& B = b->batch;
    if (n_round < 0 || frame_interval < 0) fail("run_logged: need n_round >= 0 and frame_interval >= 0");
    const bool frames = frame_interval > 0 && !b->out.empty();
    const uint64_t mc = (uint64_t)b->mc_interval, fi = (uint64_t)frame_interval;
    int done = 0;
    // the order of one round of src/main.cpp:941-964: pivot move (not at round
    // 0), frame, then thermostat and integration in the launch; replica
    // exchange where a segment of the reference loop ends (:991-995), with
    // the swap sets taken in turn (:260-263)
    // with one shared potential and fixed temperatures the exchange runs
    // inside the launch: a span between frames and pivot moves is one launch
    const bool dev_rex = b->device_rex_ok();
    if (dev_rex) {
        b->rex_alloc();
        B.rex_dev = b->rex_dev;
        B.rex_restore = [b]() { b->rex_upload(); };
    }
    struct Clear {
        Batch& B;
        ~Clear() { B.rex_dev = RexDev{}; B.rex_restore = nullptr; }
    } clear{B};
    while (done < n_round) {
        const uint64_t r = B.round_num;
        if (mc && r > 0 && r % mc == 0) b->mc_step(r);
        if (frames && r % fi == 0) b->log_frame();
        int n = n_round - done;
        if (mc) n = std::min<int64_t>(n, (int64_t)(mc - r % mc));
        if (frames) n = std::min<int64_t>(n, (int64_t)(fi - r % fi));
        if (!dev_rex && b->rex_interval > 0 && b->rex_next > r) n = std::min<int64_t>(n, (int64_t)(b->rex_next - r));
        B.run(n, 0, nullptr, 0, (int)B.round_num);
        if (dev_rex) b->rex_download();
        B.round_num += n;
        B.sync_temperature();
        done += n;
        if (!dev_rex && b->rex_interval > 0 && B.round_num == b->rex_next) {
            const int set = (int)(b->rex_attempts % (long long)b->rex_sets.size());
            const auto& S = b->rex_sets[set];
            vector<int> pr, acc(S.size());
            for (auto& p : S) { pr.push_back(p[0]); pr.push_back(p[1]); }
            if (upside_batch_replica_exchange_criterion(b, (int)S.size(), pr.data(), B.round_num, b->rex_criterion,
                                                        acc.data()))
                fail("replica exchange failed");
            for (size_t p = 0; p < S.size(); ++p) {
                b->rex_stats[set][p][1]++;
                if (acc[p]) {
                    b->rex_stats[set][p][0]++;
                    std::swap(b->replica_of[S[p][0]], b->replica_of[S[p][1]]);
                }
            }
            b->rex_attempts++;
            b->rex_next = UpsideBatch::next_swap_round(B.round_num, b->rex_interval);
        }
    }
    return 0;
    ABI_CATCH(1)
}

int upside_batch_close_output(UpsideBatch* b) {
    ABI_TRY
    for (auto& o : b->out) o->close();
    b->out.clear();
    return 0;
    ABI_CATCH(1)
}

// Loop state of the wrapper after the device state: Monte Carlo interval and
// counts, replica-exchange configuration, replica indices, pair statistics and
// next swap round, recentering mode (ints and 64-bit counts in one stream).
static vector<long long> loop_state(const UpsideBatch* b) {
    auto bits = [](double x) { long long r; std::memcpy(&r, &x, sizeof(r)); return r; };
    vector<long long> v{0x55505354415445ll, b->mc_interval, b->recenter, b->rex_interval, b->rex_criterion,
                        b->rex_attempts, (long long)b->rex_next, b->batch.anneal, bits(b->batch.anneal_factor),
                        bits(b->batch.anneal_start), bits(b->batch.anneal_end), bits(b->batch.anneal_duration),
                        (long long)b->mc_total.size()};
    for (auto& t : b->mc_total) { v.push_back(t[0]); v.push_back(t[1]); }
    v.push_back((long long)b->rex_sets.size());
    for (size_t s = 0; s < b->rex_sets.size(); ++s) {
        v.push_back((long long)b->rex_sets[s].size());
        for (size_t p = 0; p < b->rex_sets[s].size(); ++p) {
            v.push_back(b->rex_sets[s][p][0]);
            v.push_back(b->rex_sets[s][p][1]);
            v.push_back(b->rex_stats[s][p][0]);
            v.push_back(b->rex_stats[s][p][1]);
        }
    }
    v.push_back((long long)b->replica_of.size());
    for (int x : b->replica_of) v.push_back(x);
    return v;
}

long long upside_batch_state_size(UpsideBatch* b) {
    ABI_TRY
    return (long long)(b->batch.state_size() + sizeof(long long) * (1 + loop_state(b).size()));
    ABI_CATCH(-1)
}

int upside_batch_save_state(UpsideBatch* b, void* buf, long long size) {
    ABI_TRY
    if (!buf || size < upside_batch_state_size(b)) fail("save_state: buffer too small");
    b->batch.save_state((char*)buf, (size_t)size);
    const vector<long long> v = loop_state(b);
    const long long n = (long long)v.size();
    char* p = (char*)buf + b->batch.state_size();
    std::memcpy(p, &n, sizeof(n));
    std::memcpy(p + sizeof(n), v.data(), v.size() * sizeof(long long));
    return 0;
    ABI_CATCH(1)
}

int upside_batch_load_state(UpsideBatch* b, const void* buf, long long size) {
    ABI_TRY
    if (!buf || size < 0) fail("load_state: no buffer");
    b->batch.load_state((const char*)buf, (size_t)size);
    // the loop state follows the device state (whose size is now this batch's)
    const size_t at = b->batch.state_size();
    long long n = 0;
    if ((size_t)size < at + sizeof(n)) fail("load_state: truncated state");
    std::memcpy(&n, (const char*)buf + at, sizeof(n));
    if (n < 13 || (size_t)size < at + sizeof(n) * (1 + (size_t)n)) fail("load_state: truncated state");
    vector<long long> v(n);
    std::memcpy(v.data(), (const char*)buf + at + sizeof(n), (size_t)n * sizeof(long long));
    if (v[0] != 0x55505354415445ll) fail("load_state: not a saved batch state");
    size_t i = 1;
    auto next = [&]() {
        if (i >= v.size()) fail("load_state: truncated loop state");
        return v[i++];
    };
    b->mc_interval = (int)next();
    b->recenter = (int)next();
    b->rex_interval = (int)next();
    b->rex_criterion = (int)next();
    b->rex_attempts = next();
    b->rex_next = (uint64_t)next();
    auto dbl = [](long long x) { double r; std::memcpy(&r, &x, sizeof(r)); return r; };
    b->batch.anneal = (int)next();
    b->batch.anneal_factor = dbl(next());
    b->batch.anneal_start = dbl(next());
    b->batch.anneal_end = dbl(next());
    b->batch.anneal_duration = dbl(next());
    b->batch.sync_temperature();
    b->mc_total.assign((size_t)next(), {0, 0});
    for (auto& t : b->mc_total) { t[0] = next(); t[1] = next(); }
    b->mc_frame.assign(b->mc_total.size(), {0, 0});
    b->rex_sets.assign((size_t)next(), {});
    b->rex_stats.assign(b->rex_sets.size(), {});
    for (size_t s = 0; s < b->rex_sets.size(); ++s) {
        const size_t np = (size_t)next();
        for (size_t p = 0; p < np; ++p) {
            const int x = (int)next(), y = (int)next();
            const long long su = next(), at2 = next();
            b->rex_sets[s].push_back({x, y});
            b->rex_stats[s].push_back({su, at2});
        }
    }
    b->replica_of.assign((size_t)next(), 0);
    for (auto& x : b->replica_of) x = (int)next();
    const int ns = (int)b->batch.sys.size();
    b->rex_partner.assign(ns, {});
    for (size_t s = 0; s < b->rex_sets.size(); ++s)
        for (size_t p = 0; p < b->rex_sets[s].size(); ++p)
            for (int x : {b->rex_sets[s][p][0], b->rex_sets[s][p][1]})
                if (x >= 0 && x < ns) b->rex_partner[x].push_back({(int)s, (int)p});
    return 0;
    ABI_CATCH(1)
}


int upside_batch_run(UpsideBatch* b, int n_round, int frame_interval, float* frame_pot, int n_frames) {
    ABI_TRY
    if (n_round < 0) fail("run: negative round count");
    b->batch.run(n_round, frame_interval, frame_pot, n_frames, (int)b->batch.round_num);
    b->batch.round_num += n_round;
    b->batch.sync_temperature();
    return 0;
    ABI_CATCH(1)
}

int upside_batch_energies(UpsideBatch* b, float* energies) {
    ABI_TRY
    b->batch.evaluate(1);
    for (size_t k = 0; k < b->batch.sys.size(); ++k) energies[k] = b->batch.potential((int)k);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_get_deriv(UpsideBatch* b, int sys, float* deriv) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("get_deriv: system out of range");
    const Topology& T = B.sys[sys].prog->T;
    hip_check(hipMemcpy(deriv, B.sys[sys].arena + T.nodes[0].sens, (size_t)T.n_atom * 3 * sizeof(float),
                        hipMemcpyDeviceToHost), "deriv");
    return 0;
    ABI_CATCH(1)
}

int upside_batch_replica_exchange_criterion(UpsideBatch* b, int n_pairs, const int* pairs, uint64_t round,
                                            int criterion, int* accepted) {
    ABI_TRY
    if (criterion != 0 && criterion != 1) fail("replica exchange: exchange criterion must be 0 or 1");
    auto& B = b->batch;
    int ns = (int)B.sys.size();
    vector<float> beta(ns);
    for (int k = 0; k < ns; ++k) beta[k] = 1.f / B.sys[k].temperature;
    auto lboltz = [&]() {
        B.evaluate(1);
        vector<float> r(ns);
        for (int k = 0; k < ns; ++k) r[k] = -beta[k] * B.potential(k);
        return r;
    };
    auto swap_pos = [&](int s1, int s2) {
        const Topology& T1 = B.sys[s1].prog->T;
        const Topology& T2 = B.sys[s2].prog->T;
        if (T1.n_atom != T2.n_atom) fail("replica exchange between systems of different size");
        size_t bytes = (size_t)T1.n_atom * 3 * sizeof(float);
        vector<float> a(T1.n_atom * 3), c(T1.n_atom * 3);
        hip_check(hipMemcpy(a.data(), B.sys[s1].arena + T1.nodes[0].out, bytes, hipMemcpyDeviceToHost), "swap");
        hip_check(hipMemcpy(c.data(), B.sys[s2].arena + T2.nodes[0].out, bytes, hipMemcpyDeviceToHost), "swap");
        hip_check(hipMemcpy(B.sys[s1].arena + T1.nodes[0].out, c.data(), bytes, hipMemcpyHostToDevice), "swap");
        hip_check(hipMemcpy(B.sys[s2].arena + T2.nodes[0].out, a.data(), bytes, hipMemcpyHostToDevice), "swap");
    };
    for (int p = 0; p < n_pairs; ++p)
        if (pairs[2 * p] < 0 || pairs[2 * p] >= ns || pairs[2 * p + 1] < 0 || pairs[2 * p + 1] >= ns)
            fail("swap pair out of range");
    // RandomGenerator(seed, REPLICA_EXCHANGE_RANDOM_STREAM, 0, round), one uniform per test that needs it
    uint32_t key[4] = {B.base_seed, 1u, 0u, 0u};
    uint32_t ctr[4] = {(uint32_t)(round & 0xffffffffull), (uint32_t)(round >> 32), 0u, 0u};
    auto uniform = [&]() {
        uint32_t bits[4];
        threefry_host(ctr, key, bits);
        ctr[3]++;
        return u01_open_closed(bits[0]);
    };
    if (criterion == 0) {
        // full change of log Boltzmann weight: swap all, re-evaluate, undo rejected (:278-297)
        auto old_lb = lboltz();
        for (int p = 0; p < n_pairs; ++p) swap_pos(pairs[2 * p], pairs[2 * p + 1]);
        auto new_lb = lboltz();
        for (int p = 0; p < n_pairs; ++p) {
            int s1 = pairs[2 * p], s2 = pairs[2 * p + 1];
            float diff = (new_lb[s1] + new_lb[s2]) - (old_lb[s1] + old_lb[s2]);
            const bool reject = diff < 0.f && std::exp(diff) < uniform();
            if (reject) swap_pos(s1, s2);
            if (accepted) accepted[p] = reject ? 0 : 1;
        }
    } else {
        // log Boltzmann weights of the current positions only (:298-311)
        auto lb = lboltz();
        for (int p = 0; p < n_pairs; ++p) {
            int s1 = pairs[2 * p], s2 = pairs[2 * p + 1];
            float diff = lb[s2] - lb[s1];
            const bool accept = diff >= 0.f || std::exp(diff) >= uniform();
            if (accept) swap_pos(s1, s2);
            if (accepted) accepted[p] = accept ? 1 : 0;
        }
    }
    return 0;
    ABI_CATCH(1)
}

int upside_batch_replica_exchange(UpsideBatch* b, int n_pairs, const int* pairs, uint64_t round, int* accepted) {
    return upside_batch_replica_exchange_criterion(b, n_pairs, pairs, round, 0, accepted);
}

int upside_batch_potential_terms(UpsideBatch* b, int sys, float* values, int max_terms, char* names, int names_len) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("system index out of range");
    const Topology& T = B.sys[sys].prog->T;
    vector<float> p(std::max(T.n_pot, 1));
    hip_check(hipMemcpy(p.data(), B.sys[sys].arena + T.pot, p.size() * 4, hipMemcpyDeviceToHost), "potential");
    string nm;
    for (int slot = 0; slot < T.n_pot; ++slot) {
        for (auto& n : T.nodes)
            if (n.potential && n.pot_slot == slot) nm += n.name + "\n";
        if (values && slot < max_terms) values[slot] = p[slot];
    }
    if (names && names_len > 0) {
        size_t k = std::min(nm.size(), (size_t)names_len - 1);
        std::memcpy(names, nm.data(), k);
        names[k] = 0;
    }
    return T.n_pot;
    ABI_CATCH(-1)
}

int upside_batch_bp_stats(UpsideBatch* b, int sys, int* last_iters, long long* total_iters, long long* solves,
                          long long* bad_solves) {
    ABI_TRY
    auto& B = b->batch;
    const Topology& T = B.sys[sys].prog->T;
    long long tot = 0, sol = 0, bad = 0;
    int last = 0;
    for (auto& R : T.rots) {
        int st[4];
        hip_check(hipMemcpy(st, B.sys[sys].iarena + R.stats, sizeof(st), hipMemcpyDeviceToHost), "stats");
        last += st[0];
        bad += st[1];
        tot += st[2];
        sol += st[3];
    }
    if (last_iters) *last_iters = last;
    if (total_iters) *total_iters = tot;
    if (solves) *solves = sol;
    if (bad_solves) *bad_solves = bad;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_bp_lds_solves(UpsideBatch* b, int sys, long long* n) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("bp_lds_solves: system out of range");
    long long t = 0;
    for (auto& R : B.sys[sys].prog->T.rots) {
        int v = 0;
        hip_check(hipMemcpy(&v, B.sys[sys].iarena + R.stats + 6, sizeof(int), hipMemcpyDeviceToHost), "stats");
        t += v;
    }
    *n = t;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_bp_histogram(UpsideBatch* b, int sys, long long* counts, int n_bins) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("bp_histogram: system out of range");
    if (n_bins < BP_HIST_BINS) fail("bp_histogram: need " + std::to_string(BP_HIST_BINS) + " bins");
    const Topology& T = B.sys[sys].prog->T;
    for (int k = 0; k < n_bins; ++k) counts[k] = 0;
    for (auto& R : T.rots) {
        std::vector<int> h(BP_HIST_BINS);
        hip_check(hipMemcpy(h.data(), B.sys[sys].iarena + R.stats + BP_HIST, BP_HIST_BINS * sizeof(int),
                            hipMemcpyDeviceToHost), "stats");
        for (int k = 0; k < BP_HIST_BINS; ++k) counts[k] += h[k];
    }
    return BP_HIST_BINS;
    ABI_CATCH(-1)
}

float upside_batch_last_run_ms(UpsideBatch* b) { return b ? b->batch.last_ms : -1.f; }

int upside_batch_max_team(UpsideBatch* b) {
    ABI_TRY
    return b->batch.max_team();
    ABI_CATCH(-1)
}

int upside_batch_set_team(UpsideBatch* b, int team) {
    ABI_TRY
    b->batch.set_team(team);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_get_team(UpsideBatch* b) { return b ? b->batch.team : -1; }

int upside_batch_set_lane_split(UpsideBatch* b, int split) {
    ABI_TRY
    if (split != 2 && split != 4) fail("lane split must be 2 or 4");
    b->batch.lane_split = split;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_get_lane_split(UpsideBatch* b) { return b ? b->batch.lane_split : -1; }

int upside_check_program(const char* up_path, char* msg, int msg_len) {
    string out;
    int rc = 0;
    try {
        if (!up_path) return -1;
        Program P;
        P.path = up_path;
        P.load(0, 1.0, false);
        out = P.check_program();
        if (!out.empty()) {
            rc = 1;
        } else {
            int in_sections = 0;
            for (auto& sec : P.lane_sections) in_sections += sec.second - sec.first;
            out = std::to_string(P.n_germ()) + " value ops, " + std::to_string(P.n_deriv()) + " deriv ops, " +
                  std::to_string(P.lane_sections.size()) + " sections of " + std::to_string(P.n_lanes) +
                  " lanes holding " + std::to_string(in_sections) + " ops";
        }
    } catch (std::exception& e) {
        out = e.what();
        rc = -1;
    }
    if (msg && msg_len > 0) {
        std::strncpy(msg, out.c_str(), (size_t)msg_len - 1);
        msg[msg_len - 1] = 0;
    }
    return rc;
}

int upside_plan_team_map(const int* teams, int n_sys, int per_xcd, int* map, int map_len) {
    try {
        if (!teams || n_sys < 1 || per_xcd < 1) return -1;
        vector<int> m;
        if (!plan_team_map(vector<int>(teams, teams + n_sys), per_xcd, &m)) return -1;
        if (map) {
            if ((int)m.size() > map_len) return -1;
            std::copy(m.begin(), m.end(), map);
        }
        return (int)m.size();
    } catch (...) {
        return -1;
    }
}

int upside_plan_balanced_teams(const int* n_atom, int n_sys, int budget, int per_xcd, int* teams) {
    try {
        if (!n_atom || !teams || n_sys < 1 || per_xcd < 1) return 1;
        vector<int> t = plan_balanced_teams(vector<int>(n_atom, n_atom + n_sys), budget, per_xcd);
        if (t.empty()) return 1;
        std::copy(t.begin(), t.end(), teams);
        return 0;
    } catch (...) {
        return 1;
    }
}

int upside_batch_set_teams(UpsideBatch* b, const int* teams, int n_sys) {
    ABI_TRY
    if (!teams || n_sys < 0) fail("set_teams: no team sizes");
    b->batch.set_teams(vector<int>(teams, teams + n_sys));
    return 0;
    ABI_CATCH(1)
}

int upside_batch_get_teams(UpsideBatch* b, int* teams, int n_sys) {
    ABI_TRY
    auto& B = b->batch;
    if (!teams || n_sys != (int)B.sys.size()) fail("get_teams: one entry per system expected");
    for (int s = 0; s < n_sys; ++s) teams[s] = B.team_of(s);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_balanced_teams(UpsideBatch* b, int budget, int* teams, int n_sys) {
    ABI_TRY
    auto& B = b->batch;
    if (!teams || n_sys != (int)B.sys.size()) fail("balanced_teams: one entry per system expected");
    vector<int> t = B.balanced_teams(budget);
    std::copy(t.begin(), t.end(), teams);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_autotune(UpsideBatch* b, int n_round, const int* candidates, int n_candidates, float* ms_out) {
    ABI_TRY
    b->batch.autotune(n_round, candidates, n_candidates, ms_out);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_autotune_layout(UpsideBatch* b, int n_round, const int* teams, int n_teams, const int* layouts,
                                 int n_layouts, float* ms_out) {
    ABI_TRY
    if (!layouts || n_layouts < 1) throw std::runtime_error("no rotamer solve layouts to tune");
    b->batch.autotune(n_round, teams, n_teams, ms_out, layouts, n_layouts);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_bp_layout(UpsideBatch* b, int layout) {
    ABI_TRY
    if (layout < 0 || layout > 3) throw std::runtime_error("rotamer solve layout must be 0, 1, 2 or 3");
    b->batch.bp_layout = layout;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_get_bp_layout(UpsideBatch* b) { return b ? b->batch.bp_layout : -1; }

int upside_batch_set_bp_schedule(UpsideBatch* b, int mode, float residual, int budget, int klocal) {
    ABI_TRY
    auto& B = b->batch;
    if (mode < 0 || mode > 3) fail("rotamer solve schedule must be 0, 1, 2 or 3");
    if ((mode == 1 || mode == 2) && !(residual > 0.f && residual < 1.f)) fail("residual must be in (0, 1)");
    if (mode == 3 && !(residual >= 0.f && residual < 1.f)) fail("schedule 3 tolerance must be in [0, 1)");
    if (mode >= 2 && !B.mlp) fail("schedules 2 and 3 need the warm-start network (upside_batch_set_warm_start)");
    if (budget < 2 || klocal < 0 || klocal > 64) fail("budget must be at least 2 iterations and klocal in [0, 64]");
    B.bp_mode = mode;
    B.bp_residual = residual;
    B.bp_budget = budget;
    B.bp_klocal = klocal;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_set_warm_start(UpsideBatch* b, const float* params, int n_param) {
    ABI_TRY
    auto& B = b->batch;
    if (n_param != MLP_NPARAM) fail("warm-start parameters: " + std::to_string(MLP_NPARAM) + " floats expected");
    if (!B.mlp) B.mlp.reset(new MlpParams());
    B.mlp->set(params, B.stream);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_bp_training_rows(UpsideBatch* b, int sys, float* features, float* targets, int max_rows) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("bp_training_rows: system out of range");
    const Topology& T = B.sys[sys].prog->T;
    int n = 0;
    for (auto& R : T.rots) {
        const int nn = R.n_node;
        const int* nrot = T.ci.data() + R.node_nrot;
        vector<float> f((size_t)nn * MLP_IN), cur((size_t)nn * 6);
        hip_check(hipMemcpy(f.data(), B.sys[sys].arena + R.feat, f.size() * 4, hipMemcpyDeviceToHost), "rows");
        hip_check(hipMemcpy(cur.data(), B.sys[sys].arena + R.ncur, cur.size() * 4, hipMemcpyDeviceToHost), "rows");
        for (int x = 0; x < nn; ++x) {
            if (nrot[x] < 2) continue;
            if (features && targets) {
                if (n >= max_rows) fail("bp_training_rows: more rows than max_rows");
                std::memcpy(features + (size_t)n * MLP_IN, f.data() + (size_t)x * MLP_IN, MLP_IN * 4);
                for (int r = 0; r < 8; ++r) targets[(size_t)n * 8 + r] = r < 6 ? cur[(size_t)x * 6 + r] : 0.f;
            }
            ++n;
        }
    }
    return n;
    ABI_CATCH(-1)
}

int upside_batch_bp_warm_stats(UpsideBatch* b, int sys, long long* warm, long long* fallbacks,
                               long long* residual_solves) {
    ABI_TRY
    auto& B = b->batch;
    if (sys < 0 || sys >= (int)B.sys.size()) fail("bp_warm_stats: system out of range");
    long long w = 0, f = 0, r = 0;
    for (auto& R : B.sys[sys].prog->T.rots) {
        int st[10];
        hip_check(hipMemcpy(st, B.sys[sys].iarena + R.stats, sizeof(st), hipMemcpyDeviceToHost), "stats");
        w += st[7];
        f += st[8];
        r += st[9];
    }
    if (warm) *warm = w;
    if (fallbacks) *fallbacks = f;
    if (residual_solves) *residual_solves = r;
    return 0;
    ABI_CATCH(1)
}

int upside_batch_enable_profile(UpsideBatch* b, int on) {
    ABI_TRY
    b->batch.enable_profile(on != 0);
    return 0;
    ABI_CATCH(1)
}

int upside_batch_profile_report(UpsideBatch* b, int sys, char* buf, int buflen) {
    ABI_TRY
    string r = b->batch.profile_report(sys);
    if ((int)r.size() + 1 > buflen) return (int)r.size() + 1;
    std::memcpy(buf, r.c_str(), r.size() + 1);
    return 0;
    ABI_CATCH(-1)
}

}  // extern "C"
