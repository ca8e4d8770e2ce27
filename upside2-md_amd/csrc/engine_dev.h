// Device-side data model of the MI355X Upside engine.
//
// One system (a chain replica) is evaluated by one workgroup.  The node graph
// of a `.up` file (src/deriv_engine.cpp:353-426) is lowered on the host into a
// program of OpRec records; the persistent kernel interprets the program in
// germ order (values) and deriv order (sensitivity pushes), mirroring
// DerivEngine::compute (src/deriv_engine.cpp:239-260).  All node outputs,
// sensitivities and scratch live in per-system float/int arenas addressed by
// offsets, so replicas that share a topology share one program.
#pragma once
#include <cstdint>
#include "dev_mlp.h"

namespace upside {

// Dynamic LDS per workgroup (159 KB of the 160 KB a gfx950 workgroup may
// hold, the rest being the static reduction scratch): staging for pair-graph
// coefficient tables (the largest, the rotamer pair table, is 20*20*54 floats
// = 86.4 KB, plus read-past padding), cell grids, and the rotamer solve state.
constexpr int LDS_FLOATS = 40704;
// Rotamer statistics words: [0] iterations of the last solve, [1] solves at
// the iteration limit, [2] total iterations, [3] solves, [4] BP edges, [5]
// residue-pair candidates, [6] solves held in LDS, [7] warm-started solves,
// [8] warm starts that fell back to the plain start, [9] solves stopped by
// the residual rule (bp_mode 1 and 2), then a histogram of iterations per
// solve (bin = iterations, the last bin collects the rest).
constexpr int BP_HIST = 12;
constexpr int BP_HIST_BINS = 256;
constexpr int NT = 512;            // threads per workgroup (8 wave64)
constexpr int NWAVE = NT / 64;

enum OpKind : int {
    OP_DIST3D = 1, OP_ANGLE, OP_DIHEDRAL, OP_SPRING, OP_RAMA_COORD, OP_CAT_POS,
    OP_RAMA_MAP_POT, OP_AFFINE, OP_PLACEMENT, OP_INFER_HO, OP_PROTEIN_HBOND,
    OP_HBOND_COVERAGE, OP_HBOND_ENERGY, OP_ROTAMER, OP_WEIGHTED_POS, OP_ENV_COVERAGE,
    OP_HB_ENV_COVERAGE, OP_SIGMOID_COUPLING, OP_BB_SIGMOID_COUPLING, OP_BACKBONE_PAIRS,
    OP_KIND_COUNT
};

// Edge functions of the pair graphs (InteractionGraph IType's).
enum EdgeKind : int {
    EDGE_ROTAMER = 0,      // PosQuadSplineInteraction, src/bead_interaction.h:184-221
    EDGE_HBOND_COVERAGE,   // HBondCoverageInteraction, src/hbond.cpp:249-294
    EDGE_PROTEIN_HBOND,    // ProteinHBondInteraction, src/hbond.cpp:158-246
    EDGE_ENV_WEIGHTED,     // EnvironmentCoverageInteraction, src/environment.cpp:114-170
    EDGE_ENV_HBOND,        // HbondEnvironmentCoverageInteraction, src/environment.cpp:172-228
    EDGE_BACKBONE          // residue-pair list of BackbonePairs, src/backbone_steric.cpp:38-146
};

// A pair graph over two element sets (or one, symmetric).  Candidate pairs
// within the cache radius ("slots") are kept in compressed rows: row i1 holds
// its partners i2 (symmetric graphs: only i2 > i1) in a canonical order fixed
// by the positions alone.  The transposed rows list, for each i2, the slots
// naming it, in the same kind of canonical order.  Both orders are canonical, so every sum over slots (energies,
// sensitivity gathers) runs in a fixed order: bit-reproducible, no float
// atomics.  Per-slot derivatives are stored structure-of-arrays: component q
// of slot s at d1[q*ecap + s].
struct PairGraph {
    int kind;
    int sym;
    int n1, n2;            // element counts
    int dim1, dim2;        // coordinate widths read from the producers
    int w1, w2;            // producer row strides
    int out1, out2;        // producer output offsets (float arena)
    int sens1, sens2;      // producer sensitivity offsets (float arena)
    int loc1, loc2;        // const int arrays: element -> producer row
    int type1, type2;      // const int arrays
    int id1, id2;          // const int arrays
    int param;             // const float array, (n_type1*n_type2, n_param) + tail padding
    int n_type1, n_type2, n_param;
    int param_len;         // floats of the table staged in LDS (0: read from memory)
    int excl;              // pair accepted if |id1-id2| > excl; -1: rotamer rule (id>>4 differs)
    float cutoff2;         // refine threshold (float cutoff squared)
    float cache_cut2;      // cache threshold
    float cache_cut;       // cache radius (cell size of the grid)
    float drift2;          // rebuild when an element moved more than this (squared)
    int ddim1, ddim2;      // derivative widths (== dim except for backbone pairs)
    int ecap;              // slot capacity
    int gather_split;      // side-1 and side-2 sensitivities share producer rows: gather them in two phases
    // per-system scratch (arena offsets)
    int cache_pos;         // float (n1+n2)*3 (symmetric: n1*3)
    int pos_pack;          // float (n1+n2)*3 current packed coordinates of both sides
    int off;               // int n1+1  row offsets
    int nbr;               // int ecap  partner of each slot
    int row;               // int ecap  row of each slot
    int toff;              // int n2+1  transposed row offsets
    int tslot;             // int ecap  slot of each transposed entry
    int tmp;               // int ecap  transposed entries in arrival order during a rebuild
    int tcur;              // int max(n1,n2)+1  cursors during a rebuild
    int val;               // float ecap   edge value (0 when inactive)
    int d1, d2;            // float ddim1*ecap, ddim2*ecap (SoA)
    int esens;             // float ecap   edge sensitivity
    int act;               // int ecap     active flag
    int state;             // int [0]=cache valid, [1]=overflow, [2]=rebuild count, [3]=slots
};

// Parameter derivative of a spline pair graph after an evaluation: one
// workgroup per interaction-table row (type pair).
struct ParamDerivArgs {
    const float* x1;       // producer output of side 1, row stride w1, element rows loc1
    const float* x2;       // side 2 (symmetric graphs: side 1)
    int w1, w2;
    const int* loc1;
    const int* loc2;
    const int* type1;
    const int* type2;
    const int* row;        // slot -> side-1 element
    const int* nbr;        // slot -> side-2 element
    const int* act;        // slot refined in (edge present)
    const float* es;       // slot edge sensitivity
    int n_slot;
    const float* param;    // (n_type1*n_type2, n_param)
    int n_type2, n_param, sym, kind;
    float* out;            // (n_type1*n_type2, n_param)
};

// Parameter derivative of a sigmoid coupling node after an evaluation
// (one workgroup; one output entry per thread at a time).
struct CouplingDerivArgs {
    int kind;              // OP_SIGMOID_COUPLING or OP_BB_SIGMOID_COUPLING
    int n_res, nt, n_indep;
    const float* cov;      // coverage per residue and amino-acid type [n_res*nt]
    const float* cov_hb;   // backbone coupling: hbond coverage per residue [n_res]
    const int* ctype;      // sigmoid coupling: coupling type per residue
    const float* weights;  // [nt*nt] (sigmoid) or [nt] (backbone)
    const float* scale;    // [nt] (sigmoid)
    const float* center;   // [nt] (sigmoid)
    const float* sharp;    // [nt] (sigmoid)
    float bb_center, bb_sharp, bb_hbond_weight;
    // hbond energy: protein_hbond output (hbond value in column 6) and Rama
    // output rows, residue lists of hbond sites (CSR), the 12 parameters
    const float* pp;
    const float* rama;
    int wp, wr;
    const int* h1o;
    const int* h1i;
    const int* h2o;
    const int* h2i;
    float hp[12];
    int n_out;
    float* out;
};

// Rotamer belief propagation over residue nodes (src/rotamer.cpp:628-1169).
// Each residue node has one bead per rotamer state, listed in state order.
struct RotamerDev {
    int n_node;            // residues with side chains
    int node_nrot;         // const int [n_node]
    int node_bead0;        // const int [n_node+1]  beads of node n: [bead0[n], bead0[n+1])
    int bead_rot;          // const int [n_bead]
    int bead_node;         // const int [n_bead]
    int n_prob;            // number of 1-body energy producers (<=4)
    int prob_out[4], prob_sens[4], prob_w[4];
    float damping, tol;
    int max_iter, chunk;
    int ccap;              // residue-pair candidate capacity
    // per-system scratch (arena offsets)
    int coff;              // int n_node+1   candidates of first node a: [coff[a], coff[a+1])
    int ca, cb;            // int ccap       candidate residues (a < b)
    int slot_cand;         // int ecap of the bead graph: candidate of every slot
    int ioff;              // int n_node+1   incidence offsets
    int iinc;              // int 2*ccap     incidences: candidate*2 + (node is b)
    int eidx;              // int ccap+1     BP edge of a candidate (exclusive scan of presence)
    int bpc, bpf, bps;     // int ccap       BP edge: packed endpoints (first | second << 16), first node
                           //                (fewer rotamers), second node
    int ninc;              // int n_node     BP incidences of a node
    int binc;              // int 2*ccap     BP incidences (edge*2 + side), stored from ioff[x]
    int P;                 // float 36*ccap  edge probability matrix, SoA over BP edges
    int msg;               // float 12*ccap  [m1(6) m2(6)] per BP edge, SoA, updated in place
    int msg2;              // float 12*ccap  second message buffer of the team solve (passes alternate)
    int marg;              // float 36*ccap  edge marginals, SoA
    int e1x;               // float n_bead   energy of edges to 1-rotamer residues
    int ncen;              // float n_node*4 bead-sphere centre and radius of each residue
    int nprob, ncur, nold; // float n_node*6
    int noff;              // float n_node   energy offset
    int feat;              // float n_node*MLP_IN  warm-start features of the last solve (bp_mode 1, 2)
    int nprev;             // float n_node*6  node marginals of the previous solve (0 before the first)
    int stats;             // int [0] last iterations [1] bad solves [2] total iterations [3] solves
                           //     [4] BP edges [5] candidates
};

struct OpRec {
    int kind;
    int node;              // node index in the host graph (for diagnostics)
    int n_elem, width;     // output geometry (CoordNodes)
    int out, sens;         // output/sens offsets (float arena); potential nodes: -1
    int pot_slot;          // index into the per-system potential array (potential nodes)
    int in_out[4], in_sens[4], in_w[4];  // inputs (producer output, sens, row stride)
    int ia[16];            // kind-specific int parameters/arrays (const arena offsets)
    float fa[16];          // kind-specific float parameters
    int scratch;           // kind-specific float scratch offset
    int iscratch;          // kind-specific int scratch offset
    int plans[4];          // gather plans (const int arena offsets, -1: none)
    int graph;             // index into PairGraph table (-1 if none)
    int rot;               // index into RotamerDev table (-1 if none)
    int solo;              // run by one workgroup of the team (small elementwise op)
    int overlap_end;       // rotamer: germ ops [this+1, overlap_end) run beside the BP solve (0: none)
    int lane_end;          // first op of a two-lane section: germ ops [this, lane_end) run on two
                           // half-teams by lane (0: not a section start)
    int lane;              // in a section: 0 runs on ranks [0, G/2), 1 on [G/2, G)
    int early;             // deriv op that may run beside the rotamer solve, after its overlap ops
    int rot_graph_op;      // first op of a section: the rotamer op whose pair-list refresh and edge
    int rot_graph_lane;    // evaluation lane rot_graph_lane runs first (-1: none)
};

struct SysDesc {
    float* arena;          // per-system float arena
    int* iarena;           // per-system int arena
    const OpRec* ops;      // program: n_germ value ops then n_deriv push ops
    int n_germ, n_deriv;
    const PairGraph* graphs;
    const RotamerDev* rots;
    const float* cf;       // constant float arena (parameters)
    const int* ci;         // constant int arena (index arrays)
    int n_atom;
    int pos_out, pos_sens; // pos node offsets (row stride 3)
    int n_pot;             // number of potential nodes
    int pot;               // float offset of the per-node potential array [n_pot]
    int sens_zero_cnt;     // sens ranges to clear at the start of compute
    int sens_zero;         // int offset: pairs (offset, count)
    int mom;               // float offset: momentum [n_atom*3]
    // thermostat (src/thermostat.h, src/thermostat.cpp)
    uint32_t seed;
    float mom_scale, noise_scale;
    uint64_t n_invocations;
    float temp0;           // initial temperature (annealing schedule)
    int error;             // int offset of error word (then one per section sub-team)
    int n_lanes;           // lanes of the program's sections (2 or 4)
    int bar;               // int offset (16-byte aligned) of the team barrier words: [0..1] 64-bit
                           // arrival counter, [2] spin-timeout flag, [3] placement report, [4..4+MAX_TEAM)
                           // XCD of each workgroup, [4+MAX_TEAM..+2) 64-bit counter of the sub-team
                           // (ranks 1..G-1) that runs beside the rotamer solve
    int tred;              // float offset of team reduction scratch [2 * MAX_TEAM]
    unsigned long long* prof;  // optional per-op wall-clock accumulators (see PROF_*), or null
};

// Profile accumulator layout: [value op k] for k < n_germ, [deriv op k] at
// n_germ + k, then the fixed slots below (wall_clock64 ticks).
enum {
    PROF_BP = 0, PROF_REBUILD = 1, PROF_THERMO = 2, PROF_INTEGRATE = 3, PROF_PG_REBUILD = 4, PROF_PG_EDGES = 5,
    PROF_PG_GATHER = 6, PROF_ROT_FILL = 7, PROF_ROT_REBUILD = 8, PROF_RB_COUNT = 9, PROF_RB_FILL = 10,
    PROF_RB_MAP = 11, PROF_RB_GRID = 12, PROF_RB_SCAN = 13, PROF_BP_EDGE = 14, PROF_BP_NODE = 15,
    PROF_BP_SETUP = 16, PROF_BESIDE = 17, PROF_EXTRA = 18
};

// Replica exchange inside a persistent launch (the swap step of
// ReplicaExchange::attempt_swaps, src/main.cpp:243-316, between rounds): all
// systems share one potential and differ in temperature.  When the number of
// completed rounds reaches state[1], every team evaluates its potential and
// stages its positions, the launch meets at a grid barrier, one thread takes
// the Metropolis decisions of the current swap set with Threefry stream 1
// (counter = that round count), and after a second grid barrier every system
// of an accepted pair takes its partner's staged positions.
struct RexDev {
    int interval;          // rounds between attempts (0: no exchange in the launch)
    int criterion;         // --exchange-criterion 0 or 1 (src/main.cpp:278-311)
    int n_sets;
    const int* set_off;    // [n_sets+1] first pair of each set
    const int* pairs;      // [2 * pairs] system indices
    long long* state;      // [0] attempts so far, [1] round count of the next attempt
    long long* stats;      // [2 * pairs] successes, attempts of every pair
    int* replica_of;       // [n_sys] replica index held by each system
    const float* beta;     // [n_sys] 1 / temperature
    float* energy;         // [n_sys] potentials at the attempt
    float* stage;          // [n_sys * stage_stride] staged positions
    int stage_stride;
    int* partner;          // [n_sys] system whose positions to take (-1: keep)
    unsigned long long* grid_bar;  // grid barrier arrival counter (multiple of n_wg between launches)
    unsigned* grid_flag;           // its spin-timeout flag
    uint32_t seed;
    int n_wg;              // workgroups that take part (systems x team)
};

struct RunArgs {
    int mode;              // 1: potential and derivative, 0: derivative only
    int n_round;           // rounds to integrate (0: one evaluation only)
    float dt;
    int thermostat_interval;
    int first_round;       // global round number of the first round (for intervals)
    int frame_interval;    // rounds between potential evaluations recorded
    float* frame_pot;      // optional [n_sys * n_frames] potentials at frames
    int n_frames;
    int team;              // workgroups per system (1: no grid barriers)
    int n_sys;             // systems in the launch
    // per-system teams (null: every system has `team`): entry b of the
    // n_blocks-long map is (sys << 14) | (G << 7) | rank for workgroup b, -1
    // for a padding workgroup; workgroup b runs on XCD b % N_XCD, and the host
    // packs every team onto one XCD
    const int* blockmap;
    int n_blocks;
    // sections of four-lane programs on teams of eight or more workgroups:
    // 4 runs each lane on a quarter of the team, else two halves run two
    // lanes each (the same results either way: the lanes are independent)
    int lane_split;
    // simulated annealing (src/main.cpp:593-612, :975-977): when anneal is set
    // the thermostat temperature at round r is the schedule at time
    // 3 dt (r+1) from each system's initial temperature
    int anneal;
    double anneal_factor, anneal_start, anneal_end, anneal_duration;
    // rotamer solves: 0 on one workgroup when the BP state fits its LDS, else
    // on the whole team; 1 always the team; 2 always one workgroup
    int bp_layout;
    // rotamer solve schedule: 0 the reference's (start from the node
    // probabilities, stop when the one-sided change is at most the node's
    // tol); 1 the same start, stop when the two-sided change over a chunk is
    // at most bp_residual; 2 warm start (first pass, then the network's
    // beliefs and bp_klocal local message updates), the residual rule, and
    // after bp_budget iterations without meeting it a restart as schedule 1
    int bp_mode;
    float bp_residual;
    int bp_budget;
    int bp_klocal;
    MlpDev mlp;            // warm-start network (bp_mode 2)
    RexDev rex;            // replica exchange between rounds of this launch
};

constexpr int MAX_TEAM = 64;
// Team words: [0..1] the team's 64-bit arrival counter, [2] spin-timeout flag,
// [3] placement report, [4..4+MAX_TEAM) XCD of each workgroup, then the
// 64-bit arrival counters of the sub-teams: ranks 1..G-1, ranks 0..G/2-1,
// ranks G/2..G-1, and the four quarters [j*G/4, (j+1)*G/4) (each stays a
// multiple of its own size).
constexpr int BAR_WORDS = 4 + MAX_TEAM + 14;
constexpr int BAR_SUB1 = 4 + MAX_TEAM, BAR_LO = 4 + MAX_TEAM + 2, BAR_HI = 4 + MAX_TEAM + 4;
constexpr int BAR_Q = 4 + MAX_TEAM + 6;
// Lanes of a program's sections (OpRec::lane < MAX_LANES); error words
// error+1 .. error+MAX_LANES belong to the sub-teams of a section.
constexpr int MAX_LANES = 4;
// Workgroups of a launch are dispatched round-robin over this many XCDs
// (MI355X: 8 XCDs of 32 CUs, each with its own L2).  Team launches place every
// team on one XCD (see engine_kernels.hip), which the kernel verifies.
constexpr int N_XCD = 8;

}  // namespace upside

namespace upside {
// Gather plan header in the constant int arena:
//   [0] n_target  [1] cw (columns)  [2] target sens offset  [3] target stride
//   [4] target column  [5] contribution base (float arena)  [6] contribution
//   row stride  [7] n_idx, then off[n_target+1], then idx[n_idx].
// target row t receives sum_{j in [off[t],off[t+1])} contrib[idx[j]*stride + c].
enum { PLAN_HDR = 8 };
}  // namespace upside
