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
