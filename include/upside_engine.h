/* C ABI of the MI355X Upside force engine (libupside.so).
 *
 * The first group of entry points has exactly the signatures and return-code
 * conventions of the reference's engine_c_library (src/engine_c_library.h:7-32,
 * src/engine_c_library.cpp), which py/upside_engine.py binds over ctypes
 * (py/upside_engine.py:48-91): 0 is success, nonzero failure with a message on
 * stderr, no exception crosses the boundary, the caller owns every buffer.
 * Positions and derivatives are (n_atom, 3) row-major float32; evaluate_deriv
 * returns dE/dx (the force is its negative).
 *
 * The second group extends the ABI for many systems evaluated and integrated
 * in one persistent GPU launch (replica exchange, independent chains).  It has
 * no counterpart in the reference, whose only batching is an OpenMP loop over
 * systems inside upside_main (src/main.cpp:915-999).
 */
#ifndef UPSIDE_ENGINE_H
#define UPSIDE_ENGINE_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct DerivEngine DerivEngine;

/* src/engine_c_library.cpp:9-21: reads /input/potential of a .up file; NULL on failure. */
DerivEngine* construct_deriv_engine(int n_atom, const char* potential_file, bool quiet);
/* src/engine_c_library.cpp:24-26 */
void free_deriv_engine(DerivEngine* engine);
/* src/engine_c_library.cpp:30-46: total potential at pos. */
int evaluate_energy(float* energy, DerivEngine* engine, const float* pos);
/* src/engine_c_library.cpp:49-65: dE/dx at pos, (n_atom,3). */
int evaluate_deriv(float* deriv, DerivEngine* engine, const float* pos);
/* src/engine_c_library.cpp:68-110: node parameter vectors, in the reference's
 * layouts: pair-graph interaction tables, sigmoid_coupling (scale, center,
 * sharpness per type, then weights), bb_sigmoid_coupling (scale, center,
 * sharpness, hbond weight) and hbond_energy (12 parameters)
 * (src/environment.cpp:925-955, :1102-1115, src/hbond.cpp:653, :722-737).
 * set_param accepts tables only where the pair-list cutoff does not depend on
 * them (rotamer, hbond_coverage(_hydrophobe), protein_hbond; a symmetric table
 * must match its transpose, src/interaction_graph.h:385-396).
 * get_param_deriv (the reference's PARAM_DERIV build,
 * src/interaction_graph.h:406-416) returns d(potential)/d(parameters) at the
 * last evaluation for the spline pair graphs rotamer, hbond_coverage and
 * hbond_coverage_hydrophobe, zeros for the environment-coverage graphs (as the
 * reference, src/environment.cpp:164-167, :222-225), the derivatives of
 * sigmoid_coupling, bb_sigmoid_coupling (:958-997, :1118-1134) and
 * hbond_energy (src/hbond.cpp:655-720), and nonzero for nodes the reference
 * gives no derivative (protein_hbond, backbone_pairs). */
int set_param(int n_param, const float* param, DerivEngine* engine, const char* node_name);
int get_param_deriv(int n_param, float* deriv, DerivEngine* engine, const char* node_name);
int get_param(int n_param, float* param, DerivEngine* engine, const char* node_name);
/* src/engine_c_library.cpp:113-178: row-major (n_elem, elem_width) copies of a
 * coordinate node's output or sensitivity from the last evaluation; potential
 * nodes report (1,1) and their potential. */
int get_output_dims(int* n_elem, int* elem_width, DerivEngine* engine, const char* node_name);
int get_output(int n_output, float* output, DerivEngine* engine, const char* node_name);
int get_sens(int n_output, float* output, DerivEngine* engine, const char* node_name);
/* src/engine_c_library.cpp:181-195 and the node hooks of src/rotamer.cpp:736-834,
 * src/hbond.cpp:476-482.  Supported names: rotamer: "rotamer_free_energy"
 * (n_node), "rotamer_1body_energy" (n_node x energy inputs), "node_energy"
 * (n_node x 6), "edge_energy" and "edge_marginal_in_graph_order" (n_node x
 * n_node x 6 x 6, nodes numbered by class 1, 3, 6 rotamers as the reference),
 * "count_edges_by_type", "n_node", "read n_bad_solve", "read n_bad_solve and
 * reset" (values of the last evaluation), and "rotamer_potential" and
 * "rotamer_entropy" (n_node), the per-residue parts of the free energy that
 * the reference writes only to its detailed log (src/rotamer.cpp:716-724);
 * any pair-graph node: "count_edges_by_type", "n_edge", "edge_pairs"
 * (2*n_edge element indices, sorted), "edge_values" (n_edge, same order).
 * Measurement hooks with no reference counterpart: pair graphs "graph_shape"
 * (n1, dim1, n2, dim2); rotamer "iterations" (of the last solve),
 * "bp_edge_classes" (present residue-pair edges of classes 11 13 16 33 36 66),
 * "node_classes" (residues with 1, 3, 6 rotamers). */
int get_value_by_name(int n_output, float* output, DerivEngine* engine, const char* node_name,
                      const char* log_name);
/* src/engine_c_library.cpp:197-277: clamped cubic B-spline helpers. */
int clamped_spline_solve(int N, float* bspline_coeff, const float* values);
int clamped_spline_value(int N, float* result, const float* bspline_coeff, int nx, float* x);
int get_clamped_value_and_deriv(int N, float* result, const float* bspline_coeff, int nx, float* x);
int get_clamped_coeff_deriv(int N, float* result, const float* bspline_coeff, float x);

/* ---- batched extension (no reference counterpart) ---- */
typedef struct UpsideBatch UpsideBatch;

/* n_sys systems from n_sys .up paths (a path may repeat: replicas share one
 * lowered program).  Initial positions come from /input/pos of each file.
 * temperatures[n_sys]; system s uses thermostat seed base_seed+s
 * (src/main.cpp:682), timescale 5, interval rounds thermostat_interval with
 * inner step 3 and time step dt (src/main.cpp:775-803).  Momenta are
 * thermalised as the reference does at a fresh start.  NULL on failure. */
UpsideBatch* upside_batch_create(int n_sys, const char* const* up_paths, const float* temperatures,
                                 uint32_t base_seed, float dt, int thermostat_interval);
void upside_batch_free(UpsideBatch* batch);
int upside_batch_n_atom(UpsideBatch* batch, int sys);
int upside_batch_set_pos(UpsideBatch* batch, int sys, const float* pos);
int upside_batch_get_pos(UpsideBatch* batch, int sys, float* pos);
int upside_batch_get_mom(UpsideBatch* batch, int sys, float* mom);
int upside_batch_set_mom(UpsideBatch* batch, int sys, const float* mom);
/* Advance every system by n_round rounds (3 force evaluations each) in one
 * launch; when frame_interval>0 the potential is evaluated at rounds (counted
 * since the batch was created) that are multiples of it and stored in
 * frame_pot[n_sys*n_frames] if non-NULL. */
int upside_batch_run(UpsideBatch* batch, int n_round, int frame_interval, float* frame_pot, int n_frames);
/* Potential of every system at its current positions: energies[n_sys]. */
int upside_batch_energies(UpsideBatch* batch, float* energies);
/* dE/dx (n_atom, 3) of system sys at the last evaluation: upside_batch_energies,
 * or the last force evaluation of a run (which precedes that run's last
 * position update). */
int upside_batch_get_deriv(UpsideBatch* batch, int sys, float* deriv);
/* Replica exchange step (src/main.cpp:243-316, criterion 0): swap positions
 * of the pairs of swap set `set_index`, keep a swap with the Metropolis test
 * drawn from Threefry stream 1 with counter `round`.  accepted[n_pairs] out. */
int upside_batch_replica_exchange(UpsideBatch* batch, int n_pairs, const int* pairs, uint64_t round,
                                  int* accepted);
/* The same with the reference's --exchange-criterion (src/main.cpp:278-311):
 * 0 as above; 1 compares the log Boltzmann weights at the current positions,
 * exp(-b2 E2 + b1 E1), and swaps when the test passes. */
int upside_batch_replica_exchange_criterion(UpsideBatch* batch, int n_pairs, const int* pairs, uint64_t round,
                                            int criterion, int* accepted);
/* Per-term potentials of system `sys` at the last evaluation (each potential
 * node's `potential`, summed into the total in src/deriv_engine.cpp:240-250): values[slot] for
 * slot < max_terms, node names '\n'-separated into names.  Returns the number
 * of terms (-1 on error). */
int upside_batch_potential_terms(UpsideBatch* batch, int sys, float* values, int max_terms, char* names,
                                 int names_len);
/* Summed BP statistics: iterations of the last solve and since creation. */
int upside_batch_bp_stats(UpsideBatch* batch, int sys, int* last_iters, long long* total_iters,
                          long long* solves, long long* bad_solves);
/* Histogram of BP iterations per solve since creation, summed over the
 * system's rotamer nodes: counts[k] solves of k iterations (the last bin
 * collects the rest; the iteration count is the `iter` that
 * solve_for_marginals returns, src/rotamer.cpp:1137-1161).  n_bins must be at
 * least 256; returns the number of bins written, -1 on error.  No counterpart in the reference's C ABI (its solver keeps
 * only n_bad_solve, src/rotamer.cpp:845-846). */
int upside_batch_bp_histogram(UpsideBatch* batch, int sys, long long* counts, int n_bins);
/* Solves since creation whose BP state was held in LDS (the rest ran from
 * memory); no counterpart in the reference. */
int upside_batch_bp_lds_solves(UpsideBatch* batch, int sys, long long* n);
/* /output logging as the reference's simulation loop does at its basic log
 * level (src/main.cpp:705-736, :807-831, :893, :946-949; datasets as
 * src/state_logger.h:21-39 creates them): per system, an `output` group with
 * extendable datasets pos (n_frame,1,n_atom,3) f4, mom (same, if
 * record_momentum), kinetic (n_frame,1) f8, potential (n_frame,1) f8, time
 * (n_frame,) f8 and temperature (n_frame,1) f8, chunked by 100 frames.
 * output_paths NULL writes into each system's input .up (replacing an existing
 * /output, as the reference does without -o; systems must not share an input
 * file), else output_paths[s] is created afresh (as with -o).  log_level 1
 * (the reference's default, "detailed") adds hbond (n_frame, n_site) f4, the
 * hbond value of every donor and acceptor (src/hbond.cpp:373-377),
 * rotamer_bad_solves_cumulative (n_frame, 1) i8 (src/rotamer.cpp:708-710), the
 * per-residue rotamer_free_energy, rotamer_potential, rotamer_entropy and
 * rotamer_1body_energy<k> (src/rotamer.cpp:711-731), rama_map_potential
 * (n_frame, n_res) of a rama_map_pot node with log_pot set
 * (src/rama_map_pot.cpp:49-55) and sigmoid_coupling (n_frame, n_res)
 * (src/environment.cpp:892-900); log_level 2 ("extensive") also
 * environment_coverage and hbond_environment_coverage (n_frame, n_elem), the
 * coverage node outputs (src/environment.cpp:304-310, :355-360), and virtual
 * (n_frame, n_site, 3), the inferred H and O positions (src/hbond.cpp:53-58).
 * Where two nodes would log under one name (the reference would fail to
 * create the second dataset) the first in load order is logged.
 * upside_batch_run_logged advances n_round rounds and records a frame at
 * every round number (counted since the batch was created) that is a
 * multiple of frame_interval (> 0), before that round's thermostat and
 * integration.
 * upside_batch_close_output flushes and closes the files (also done by
 * upside_batch_free).  0 on success. */
int upside_batch_open_output(UpsideBatch* batch, const char* const* output_paths, int record_momentum,
                             int log_level);
int upside_batch_run_logged(UpsideBatch* batch, int n_round, int frame_interval);
int upside_batch_close_output(UpsideBatch* batch);
/* Monte Carlo pivot moves (src/monte_carlo_sampler.cpp:19-137, :255-283, from
 * each file's /input/pivot_moves, py/upside_config.py:2543-2552): with
 * mc_interval > 0, upside_batch_run_logged makes one pivot move per system at
 * every round number r > 0 that is a multiple of mc_interval, before that
 * round's frame, thermostat and integration (src/main.cpp:941-944): the move
 * is drawn from Threefry stream 2 with counter r and kept by the Metropolis
 * test on the potentials evaluated before and after.  frame_interval 0 (or no
 * output open) runs without frames; an open output then also records
 * pivot_stats (accepted, attempted since the last frame) when the interval was
 * set before upside_batch_open_output.  Fails when a file has no pivot moves.
 * upside_batch_mc_stats: accepted and attempted moves since the interval was
 * set.  No counterpart in the reference's C ABI (its CLI option
 * --monte-carlo-interval, src/main.cpp:471, :570-571). */
int upside_batch_set_monte_carlo(UpsideBatch* batch, int mc_interval);
/* Thermostat timescale (--thermostat-timescale, src/main.cpp:477, :775-781;
 * 5 by default): only before the first round; the fresh-start
 * thermalisation is redone with the new parameters. */
int upside_batch_set_thermostat_timescale(UpsideBatch* batch, float timescale);
/* Simulated annealing as --anneal-factor, --anneal-start, --anneal-end and
 * --anneal-duration (src/main.cpp:593-612, :975-977), times in the CLI's
 * units: each thermostat application at round r uses the schedule's
 * temperature at time 3 dt (r+1), which replica exchange and the logged
 * temperature then see.  The CLI's default duration (the whole run) must be
 * given as `duration` or through an end after the start; factor 1 disables. */
int upside_batch_set_annealing(UpsideBatch* batch, double factor, double start, double end, double duration);
/* Recentering before every logged frame, as the reference CLI does unless
 * --disable-recentering (src/main.cpp:576-577, :947; src/deriv_engine.cpp:39-49):
 * mode 0 none (the default here), 1 subtract the centroid, 2 its x and y only
 * (--disable-z-recentering). */
int upside_batch_set_recentering(UpsideBatch* batch, int mode);
/* Replica exchange inside upside_batch_run_logged, as the reference's system
 * loop does it with --replica-interval, --swap-set and --exchange-criterion
 * (src/main.cpp:159-316, :915-999): n_sets swap sets of set_sizes[s] pairs
 * (pairs flattened), tried in turn; a swap is attempted when the batch's round
 * counter reaches the end of a segment of the reference loop (the first
 * multiple of replica_interval at least two rounds after the previous swap),
 * with Threefry stream 1 and counter = that round.  An open output then also
 * records replica_index, replica_swap_partner and replica_cumulative_swaps
 * (:210-232) when this was set before upside_batch_open_output.
 * replica_interval 0 disables.  upside_batch_replica_stats: the replica index
 * held by each system and successes/attempts of every pair (sets in order). */
int upside_batch_set_replica_exchange(UpsideBatch* batch, int replica_interval, int criterion, int n_sets,
                                      const int* set_sizes, const int* pairs);
int upside_batch_replica_stats(UpsideBatch* batch, int* replica_index, long long* successes, long long* attempts);
int upside_batch_mc_stats(UpsideBatch* batch, int sys, long long* accepted, long long* attempted);
/* Checkpoint and resume (no counterpart in the reference, whose thermostat
 * counter lives only in memory, src/thermostat.h:24-29): the whole device state
 * of the batch (per system: thermostat seed and counter, temperature,
 * positions, momenta, Verlet caches, pair lists and BP state) into a caller
 * buffer of upside_batch_state_size() bytes.  Loading it into a batch made
 * from the same .up files (any seed, any team size) continues bit for bit as
 * the saved batch would have, including the run loop's state (Monte Carlo
 * interval and counts, replica-exchange sets, indices and counts, recentering); a state whose system count or arena sizes do not
 * match the batch is refused.  0 on success. */
long long upside_batch_state_size(UpsideBatch* batch);
int upside_batch_save_state(UpsideBatch* batch, void* buf, long long size);
int upside_batch_load_state(UpsideBatch* batch, const void* buf, long long size);
/* Synchronize and time the device (for benchmarks): elapsed ms of the last run. */
float upside_batch_last_run_ms(UpsideBatch* batch);
/* Workgroups cooperating on each system ("team"; 1 = one workgroup per
 * system, no grid barriers).  max_team is the largest team every system can
 * have with all workgroups resident at once.  autotune runs n_round rounds per
 * candidate team size from the same state, restores the state and keeps the
 * fastest; ms_out[n_candidates] (may be NULL) receives the kernel times.  A
 * candidate 0 stands for the balanced per-system teams below.
 * No reference counterpart (the reference runs each system on one thread,
 * src/main.cpp:917): set_teams gives every system its own team size
 * (teams[n_sys]; each team stays on one XCD, and all must be resident at
 * once, else an error); get_team then reports the largest.  get_teams fills
 * teams[n_sys] with the current sizes.  balanced_teams proposes sizes in
 * proportion to the systems' atom counts within `budget` workgroups (<= 0:
 * the resident maximum), without applying them.  Results do not depend on
 * the team sizes.  0 on success. */
int upside_batch_max_team(UpsideBatch* batch);
int upside_batch_set_team(UpsideBatch* batch, int team);
int upside_batch_get_team(UpsideBatch* batch);
int upside_batch_autotune(UpsideBatch* batch, int n_round, const int* candidates, int n_candidates, float* ms_out);
int upside_batch_set_teams(UpsideBatch* batch, const int* teams, int n_sys);
int upside_batch_get_teams(UpsideBatch* batch, int* teams, int n_sys);
int upside_batch_balanced_teams(UpsideBatch* batch, int budget, int* teams, int n_sys);
/* The host-side plans behind set_teams and balanced_teams, for a device with
 * per_xcd resident workgroups on each of its 8 XCDs (no reference
 * counterpart; no device calls).  upside_plan_team_map writes the block map,
 * one word per workgroup b (which runs on XCD b % 8): (system << 14) |
 * (team size << 7) | rank, or -1 for padding; it returns the map's length,
 * or -1 if the teams do not fit or map_len is too small (map may be NULL to
 * ask for the length).  upside_plan_balanced_teams fills teams[n_sys] from
 * the atom counts within budget workgroups; 0 on success. */
/* Lowers the .up file's node graph on the host (no device calls) and checks
 * the invariants the kernel relies on: the value pass is a topological order,
 * ops in different lanes of a section are independent (no dependency, no
 * shared pair graph, no potentials pushing into a shared input), the
 * rotamer's pair-list refresh and overlap ops read nothing they must wait
 * for, and the deriv pass runs each node after its consumers.  0 when all
 * hold (msg: a summary), 1 on a violation (msg: the first one), -1 when the
 * file cannot be lowered.  No reference counterpart (the scheduling it
 * checks replaces src/deriv_engine.cpp:127-178). */
int upside_check_program(const char* up_path, char* msg, int msg_len);
int upside_plan_team_map(const int* teams, int n_sys, int per_xcd, int* map, int map_len);
int upside_plan_balanced_teams(const int* n_atom, int n_sys, int budget, int per_xcd, int* teams);
/* Independent ops of a step run side by side in sections of four lanes (no
 * reference counterpart).  On a team of eight or more workgroups, split 4
 * runs each lane on a quarter of the team and split 2 (default) runs two
 * lanes on each half, one after the other; the results are the same.  The
 * autotune calls try both after choosing the team and layout.  0 on success;
 * get returns the split. */
int upside_batch_set_lane_split(UpsideBatch* batch, int split);
int upside_batch_get_lane_split(UpsideBatch* batch);
/* Where each rotamer BP solve runs (no reference counterpart; the reference
 * solves on its one thread, src/rotamer.cpp:1106-1162): 0 (default) on one
 * workgroup with the BP state in its LDS when it fits (the rest of the team
 * meanwhile evaluates the ops that do not depend on the solve), else on
 * several workgroups with the state in memory (with four or more workgroups
 * per system the first half of the team solves while the second half
 * evaluates those ops); 1 always on the whole team; 2 always on one workgroup
 * (LDS when it fits, else memory); 3 always on several workgroups as in 0.
 * Every layout computes the same arithmetic in the same order: results are
 * identical.
 * upside_batch_autotune_layout times every (team, layout) pair from the same
 * state as upside_batch_autotune does and keeps the fastest;
 * ms_out[n_teams * n_layouts] (may be NULL), row-major by team. */
int upside_batch_set_bp_layout(UpsideBatch* batch, int layout);
int upside_batch_get_bp_layout(UpsideBatch* batch);
int upside_batch_autotune_layout(UpsideBatch* batch, int n_round, const int* teams, int n_teams, const int* layouts,
                                 int n_layouts, float* ms_out);
/* Per-node wall-clock accumulation inside the kernel (the analogue of the
 * reference's -DCOLLECT_PROFILE timers, src/timing.cpp:11-54).  The report is
 * a text table; returns the needed size if buflen is too small. */
int upside_batch_enable_profile(UpsideBatch* batch, int on);
int upside_batch_profile_report(UpsideBatch* batch, int sys, char* buf, int buflen);

/* ---- BP warm-start network (no reference counterpart) ----
 * The reference starts every rotamer BP solve from the node probabilities
 * with all messages 1 (src/rotamer.cpp:1110-1133).  The warm start replaces
 * the beliefs after the first (undamped) pass by the prediction of a small
 * network: 32 features per residue -> 128 -> 128 -> 16 (ReLU), bf16 weights
 * and activations, float32 accumulation on the matrix cores.  Parameters are
 * a float32 vector of upside_mlp_n_param() entries: W1 [128][32], b1 [128],
 * W2 [128][128], b2 [128], W3 [16][128], b3 [16] (weights [out][in]).
 * Training rows are features [n][32] and targets [n][8] (the converged BP
 * marginal of each state, 0 beyond the residue's states), as
 * upside_batch_bp_training_rows returns them; the loss is the mean over rows
 * of -sum_r t_r log softmax(f[6+r] + out_r) over the residue's states.
 * upside_mlp_create: seeded He-normal hidden layers, zero output layer (the
 * untrained network leaves the first-pass beliefs unchanged).
 * upside_mlp_loss_grad: loss and gradient over all rows at the current
 * parameters (no update).  upside_mlp_train: n_steps Adam steps (lr; beta
 * 0.9, 0.999; eps 1e-8) on consecutive windows of batch_rows rows (wrapping),
 * losses[n_steps] (may be NULL) the loss of each step's batch before the
 * step.  upside_mlp_forward: the outputs [n][16] of given rows.  0 on
 * success. */
/* Rotamer BP schedule of a batch: mode 0 (default) the reference's; 1 start
 * as the reference, stop when the largest two-sided change of a node belief
 * over an iteration chunk is at most `residual`; 2 warm start: after the
 * first (undamped) pass the network's beliefs replace the pass's, klocal
 * local updates per edge make the messages consistent with them, the
 * residual rule stops the solve, and a solve that has not met it after
 * `budget` iterations restarts from the plain start as mode 1 (the fallback;
 * its iterations count too); 3 the warm start of mode 2 stopped by the
 * reference's own rule (the largest one-sided change of a belief at most the
 * rotamer node's tolerance as in mode 0, or at most `residual` when that is
 * positive), with the same budget and plain restart.  Modes 2 and 3 need
 * upside_batch_set_warm_start (parameters as upside_mlp_get_params gives
 * them).
 * upside_batch_bp_training_rows: the features of the first pass and the
 * converged node marginals of system sys's last solve in mode 1 or 2, one
 * row per residue with more than one rotamer state; with features or targets
 * NULL it returns the row count, else fills up to max_rows rows (more rows:
 * error) and returns the count; -1 on error.
 * upside_batch_bp_warm_stats: warm-started solves, fallbacks and solves that
 * met the residual rule since the batch was created. */
int upside_batch_set_bp_schedule(UpsideBatch* batch, int mode, float residual, int budget, int klocal);
int upside_batch_set_warm_start(UpsideBatch* batch, const float* params, int n_param);
int upside_batch_bp_training_rows(UpsideBatch* batch, int sys, float* features, float* targets, int max_rows);
int upside_batch_bp_warm_stats(UpsideBatch* batch, int sys, long long* warm, long long* fallbacks,
                               long long* residual_solves);

typedef struct UpsideMlp UpsideMlp;
UpsideMlp* upside_mlp_create(uint32_t seed);
void upside_mlp_free(UpsideMlp* net);
int upside_mlp_n_param(void);
int upside_mlp_get_params(UpsideMlp* net, float* params, int n_param);
int upside_mlp_set_params(UpsideMlp* net, const float* params, int n_param);
int upside_mlp_set_data(UpsideMlp* net, const float* features, const float* targets, int n_rows);
int upside_mlp_loss_grad(UpsideMlp* net, float* loss, float* grad);
int upside_mlp_train(UpsideMlp* net, int n_steps, int batch_rows, float lr, float* losses);
int upside_mlp_forward(UpsideMlp* net, const float* features, int n_rows, float* out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * allocation).  The warm-start network's loss and gradient as
 * upside_mlp_loss_grad computes them (same kernels and order, so the same
 * bits), on buffers the caller owns in device memory: params[n_param],
 * features[n_rows][32], targets[n_rows][8], out[n_param + 1] (gradient, then
 * the loss).  Every intermediate lives in a static device array of the
 * library (about 3 MB per GPU, held until the process ends); n_rows must be
 * in [1, 2^24].  Calls that overlap on one GPU share that array and race.
 * Runs on the null stream and returns when out is written; 0 on success. */
int upside_mlp_loss_grad_noalloc(const float* params, const float* features, const float* targets, int n_rows,
                                 float* out);

#ifdef __cplusplus
}
#endif
#endif
