// Test infrastructure only: linked into the oracle build of the reference
// engine (oracle/_ref), never into the product library.
//
// The reference's coord_map.cpp cannot be compiled in this image because it
// includes <Eigen/Dense> for an unrelated node (conv1d) and Eigen is absent.
// Two of its node types are part of the ff_2.1 graph, both pure index copies:
//   rama_coord (src/coord_map.cpp:69-110): out(e) = (phi[index_pos1[e]], psi[index_pos2[e]])
//   cat_pos    (src/coord_map.cpp:112-161): out = concat(pos1[index_pos1], pos2[index_pos2])
// This file registers project-written equivalents through the reference's own
// node registry (src/deriv_engine.h:219-256) so the rest of the reference graph
// can be constructed from unmodified reference sources.
#include "deriv_engine.h"
#include "h5_support.h"
#include <vector>

using namespace h5;

namespace {

struct GlueRamaCoord : public CoordNode {
    CoordNode& phi;
    CoordNode& psi;
    std::vector<int> i_phi, i_psi;

    GlueRamaCoord(hid_t grp, CoordNode& phi_, CoordNode& psi_)
        : CoordNode(get_dset_size(1, grp, "index_pos1")[0], 2), phi(phi_), psi(psi_) {
        check_size(grp, "index_pos2", n_elem);
        traverse_dset<1, int>(grp, "index_pos1", [&](size_t, int x) { i_phi.push_back(x); });
        traverse_dset<1, int>(grp, "index_pos2", [&](size_t, int x) { i_psi.push_back(x); });
    }
    void compute_value(ComputeMode) override {
        for (int e = 0; e < n_elem; ++e) {
            output(0, e) = phi.output(0, i_phi[e]);
            output(1, e) = psi.output(0, i_psi[e]);
        }
    }
    void propagate_deriv() override {
        for (int e = 0; e < n_elem; ++e) {
            phi.sens(0, i_phi[e]) += sens(0, e);
            psi.sens(0, i_psi[e]) += sens(1, e);
        }
    }
};

struct GlueCatPos : public CoordNode {
    CoordNode& a;
    CoordNode& b;
    int n_a, width;
    std::vector<int> idx;

    static int count(hid_t grp) {
        return get_dset_size(1, grp, "index_pos1")[0] + get_dset_size(1, grp, "index_pos2")[0];
    }
    GlueCatPos(hid_t grp, CoordNode& a_, CoordNode& b_)
        : CoordNode(count(grp), read_attribute<int>(grp, ".", "n_dim")), a(a_), b(b_),
          n_a(get_dset_size(1, grp, "index_pos1")[0]), width(read_attribute<int>(grp, ".", "n_dim")) {
        traverse_dset<1, int>(grp, "index_pos1", [&](size_t, int x) { idx.push_back(x); });
        traverse_dset<1, int>(grp, "index_pos2", [&](size_t, int x) { idx.push_back(x); });
    }
    void compute_value(ComputeMode) override {
        for (int e = 0; e < n_elem; ++e) {
            CoordNode& src = e < n_a ? a : b;
            for (int d = 0; d < width; ++d) output(d, e) = src.output(d, idx[e]);
        }
    }
    void propagate_deriv() override {
        for (int e = 0; e < n_elem; ++e) {
            CoordNode& src = e < n_a ? a : b;
            for (int d = 0; d < width; ++d) src.sens(d, idx[e]) += sens(d, e);
        }
    }
};

RegisterNodeType<GlueRamaCoord, 2> glue_rama("rama_coord");
RegisterNodeType<GlueCatPos, 2> glue_cat("cat_pos");

}  // namespace

// Known-answer probes of the reference's thermostat random numbers
// (src/random.h:19-67 over the vendored Random123 threefry and boxmuller),
// called by the tests to pin the oracle's restatement bit for bit.
#include "random.h"

extern "C" void ref_threefry4x32_20(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    threefry4x32_ctr_t c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    threefry4x32_key_t k = {{key[0], key[1], key[2], key[3]}};
    threefry4x32_ctr_t r = threefry4x32(c, k);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

extern "C" void ref_normal3(uint32_t seed, uint32_t stream, int n_atom, uint64_t step, float* out) {
    for (int a = 0; a < n_atom; ++a) {
        RandomGenerator g(seed, stream, (uint32_t)a, step);
        float3 v = g.normal3();
        out[a * 3 + 0] = v.x();
        out[a * 3 + 1] = v.y();
        out[a * 3 + 2] = v.z();
    }
}

extern "C" float ref_uniform_x(uint32_t seed, uint32_t stream, uint32_t index, uint64_t step) {
    RandomGenerator g(seed, stream, index, step);
    return g.uniform_open_closed().x();
}
