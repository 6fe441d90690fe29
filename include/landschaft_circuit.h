/* landschaft_circuit.h - C-ABI of the MI355X circuit library (liblandschaft_circuit.so).
 *
 * The entry points replace the reference's pybind11 bindings of its circuit
 * core (landschaft._core / landschaft._circuit_avx2, registered in
 * src/cpp/circuit/python.cpp:527-559 and reached through
 * src/landschaft/circuit/_backend.py:12-23). Arguments keep the binding's
 * names, order and meaning; arrays are plain pointers to row-major
 * (ny x nx) float64 rasters owned by the caller, outputs are written into
 * caller-allocated buffers of known size. Every call returns a status:
 *   0  success,
 *   1  invalid argument (the binding's ValueError),
 *   2  runtime failure (the binding's RuntimeError: a circuit that is not
 *      positive definite, a HIP error),
 * and lsc_last_error() gives the message of the calling thread's last failure.
 * Computation runs on the current HIP device. No entry point touches the
 * CPU solver: without a GPU, calls fail with status 2.
 */
#ifndef LANDSCHAFT_CIRCUIT_H
#define LANDSCHAFT_CIRCUIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LSC_PRECISION_MIXED = 0, LSC_PRECISION_FP64 = 1, LSC_PRECISION_AUTO = 2 };

/* Solver options (NULL: the defaults). precision: LSC_PRECISION_MIXED (FP32
 * MFMA fronts + FP64 iterative refinement, default), LSC_PRECISION_FP64, or
 * LSC_PRECISION_AUTO: Omniscape windows mixed; a graph factorisation FP64
 * when it serves at least 8 right-hand sides (pairwise jobs, Schur columns,
 * a grid circuit kept for repeated solves), mixed otherwise;
 * tol: relative residual ||b - Ax|| / ||b|| at which refinement stops
 * (default 1e-12); max_refine: refinement steps per circuit (default 30);
 * batch: circuits per batch (default 16384, fewer when HBM is short);
 * timing: 1 to time every phase with HIP events (adds synchronisations);
 * stream: a hipStream_t, 0 for the null stream. */
typedef struct lsc_solve_options {
    int32_t precision;
    int32_t max_refine;
    double tol;
    int64_t batch;
    int32_t timing;
    int32_t reserved;
    void* stream;
} lsc_solve_options;

/* One condition of Omniscape's conditional connectivity (reference
 * src/cpp/circuit/omniscape.hpp:28-33): a source cell keeps its amps when its
 * present value v matches the target block's value b of the future map
 * (median for "within", mode for "equal"): "within" lower <= v - b <= upper,
 * "equal" v == b. NaN present values always keep theirs. */
typedef struct lsc_omniscape_condition {
    const double* present; /* ny x nx; NULL: no condition */
    const double* future;  /* ny x nx; NULL: the present map */
    int32_t equal;         /* 1: "equal", 0: "within" */
    int32_t reserved;
    double lower, upper;
} lsc_omniscape_condition;

/* Per-run statistics: targets, circuits solved, refinement steps, the
 * largest and mean relative FP64 residual of the solved circuits, and (with
 * timing) milliseconds spent in each phase. The graph entry points (advanced,
 * pairwise, multi-advanced, grid circuits) always report ms_fill (host
 * symbolic analysis and upload), ms_factor and ms_solve (device). */
typedef struct lsc_solve_stats {
    int64_t n_targets;
    int64_t n_solved;
    int64_t n_unconverged; /* circuits whose residual stayed above tol */
    int64_t n_over_1e10;   /* circuits whose residual stayed above 1e-10 */
    int32_t refine_max;
    int32_t n_batches;        /* batches (launches of the factor kernel) */
    int64_t n_solve_launches; /* launches of the triangular-solve kernel */
    int64_t n_factored;       /* circuits factorised */
    double refine_mean;
    double resid_max;
    double resid_mean;
    double ms_fill, ms_flood, ms_factor, ms_solve, ms_residual, ms_currents, ms_accumulate;
    int32_t n_streams; /* Omniscape: batches in flight at once, each on its own stream (the
                          ms_ phases then add up the streams' overlapping times) */
    int32_t n_fallback; /* mixed precision: circuits (Omniscape windows) or factorisations (graphs)
                           solved again in FP64 because the FP32 factor met a non-positive pivot or
                           its refinement stalled more than 1000 times above tol */
    double resid_median; /* median relative residual of the solved circuits */
} lsc_solve_stats;

/* Message of the calling thread's last failure ("" if none). */
const char* lsc_last_error(void);

/* Device name, compute units and HBM bytes of the current device; 0 or 2. */
int lsc_device_info(char* name, int64_t name_len, int32_t* compute_units, int64_t* hbm_bytes);

/* Omniscape's cumulative current (and flow potential) over moving windows.
 * Replaces circuit_omniscape (src/cpp/circuit/python.cpp:90-137 ->
 * circuit::omniscape, src/cpp/circuit/omniscape.cpp:394-608). resistance:
 * NaN = NoData, values > 0; source: NaN = 0. conditions: NULL or two
 * conditions (the binding's condition1 / condition2, present = NULL for
 * none). current (and flow_potential_out when flow_potential != 0) receive
 * ny x nx float64 maps; n_targets the number of targets. Host pointers. */
int lsc_circuit_omniscape(const double* resistance, const double* source, int64_t nx, int64_t ny, int radius,
                          int buffer, int block_size, int four_neighbors, int resistance_is_conductance,
                          int correct_artifacts, int flow_potential,
                          const lsc_omniscape_condition* conditions, const lsc_solve_options* options,
                          double* current, double* flow_potential_out, int64_t* n_targets, lsc_solve_stats* stats);

/* The same with every array in device memory (HBM-resident inputs), the
 * condition maps included. */
int lsc_circuit_omniscape_device(const double* resistance, const double* source, int64_t nx, int64_t ny, int radius,
                                 int buffer, int block_size, int four_neighbors, int resistance_is_conductance,
                                 int correct_artifacts, int flow_potential,
                                 const lsc_omniscape_condition* conditions, const lsc_solve_options* options,
                                 double* current, double* flow_potential_out, int64_t* n_targets,
                                 lsc_solve_stats* stats);

/* The same as lsc_circuit_omniscape_device, allocating nothing: no
 * reference counterpart; it is for callers that cannot allocate device
 * memory (tight serving loops, code that must not call hipMalloc). Every
 * buffer of the run (plans, batch buffers, the correction's circuits) comes
 * from a static scratch of 128 MiB that each GPU the library runs on holds
 * until the process ends; batches are sized to fit it (one batch in flight,
 * on the caller's stream, kernels launched one by one). A run that needs
 * more than the scratch holds (one window at most, its plan, and room for
 * the FP64 fallback) fails with status 1 before any work. Calls on one GPU
 * share its scratch: the library runs them one at a time, and work that
 * overlaps a call on the same GPU must not touch its buffers. Every window is
 * solved as lsc_circuit_omniscape_device solves it; the cumulative maps
 * differ by the rounding of their sums, which add the windows of smaller
 * batches at a time. */
int lsc_circuit_omniscape_device_static(const double* resistance, const double* source, int64_t nx, int64_t ny,
                                        int radius, int buffer, int block_size, int four_neighbors,
                                        int resistance_is_conductance, int correct_artifacts, int flow_potential,
                                        const lsc_omniscape_condition* conditions, const lsc_solve_options* options,
                                        double* current, double* flow_potential_out, int64_t* n_targets,
                                        lsc_solve_stats* stats);

/* Omniscape's block-artifact correction factors, (2 (radius + buffer) + 1)^2
 * float64 into out. Replaces circuit_omniscape_correction
 * (src/cpp/circuit/python.cpp:139-153 -> omniscape.cpp:326-392). */
int lsc_circuit_omniscape_correction(int radius, int buffer, int block_size, int four_neighbors, int64_t nx,
                                     int64_t ny, const lsc_solve_options* options, double* out);

/* Circuitscape's advanced mode on a raster (sources, grounds: conductance to
 * ground, +inf = direct). Replaces circuit_advanced (src/cpp/circuit/
 * python.cpp:65-79 -> circuit::advanced, src/cpp/circuit/grid.cpp:73-210):
 * current and voltage receive ny x nx maps (voltage NaN off the nodes). */
int lsc_circuit_advanced(const double* conductance, const double* source, const double* ground, int64_t nx,
                         int64_t ny, int four_neighbors, const lsc_solve_options* options, double* current,
                         double* voltage, lsc_solve_stats* stats);

/* Advanced mode on the raster node graph (short-circuit regions merged,
 * average conductance or resistance, conflict policy 0 keepall, 1 rmvsrc,
 * 2 rmvgnd, 3 rmvall). Replaces circuitscape_advanced (src/cpp/circuit/
 * python.cpp:170-204 -> raster_graph + raster_advanced, circuitscape.cpp:
 * 145-263, 393-468). polymap may be NULL. nodemap: ny x nx; voltage and
 * current: per node (capacity nx * ny), n_nodes receives their count. */
int lsc_circuitscape_advanced(const double* gmap, const int64_t* polymap, const double* source, const double* ground,
                              int64_t nx, int64_t ny, int avg_resistances, int four_neighbors, int policy,
                              const lsc_solve_options* options, int32_t* nodemap, double* voltage, double* current,
                              int64_t* n_nodes, int32_t* solved, int32_t* n_components, lsc_solve_stats* stats);

/* Receives a pair's maps (ny x nx; NULL when not asked for) during a pairwise run. */
typedef void (*lsc_emit_fn)(int ci, int cj, const double* voltage, const double* current, void* ctx);

/* Pairwise mode: effective resistances between focal points (their cells;
 * -1: none), excluded pairs (npts x npts uint8), maps by `flags` bits
 * 1 voltage maps, 2 current maps, 4 cumulative map, 8 max map, 16 focal
 * currents zero, 32 null currents to nodata, 64 null voltages to nodata,
 * 128 log10 current maps. Replaces circuitscape_pairwise (src/cpp/circuit/
 * python.cpp:214-277 -> raster_pairwise, circuitscape.cpp:474-743, with its
 * shortcut). resistances: npts x npts; cum / max: ny x nx or NULL. */
int lsc_circuitscape_pairwise(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                              const int64_t* point_cells, int64_t npts, const uint8_t* excluded, int avg_resistances,
                              int four_neighbors, int flags, lsc_emit_fn emit, void* ctx,
                              const lsc_solve_options* options, double* resistances, double* cum, double* max,
                              lsc_solve_stats* stats);

/* Advanced mode on a network (Circuitscape's data_type = network): edges
 * a[k] - b[k] (0-based nodes < n_nodes) of conductance w[k], duplicates
 * summed, zero sums and self loops dropped; source and ground per node
 * (ground: conductance, +inf = direct); conflict policy as above.
 * Replaces network_advanced (src/cpp/circuit/python.cpp:310-346 ->
 * network_graph + raster_advanced + network_currents, circuitscape.cpp:
 * 265-362, 393-468). voltage and current: n_nodes; the branch currents
 * (nodes r < c, current) go to arrays of capacity n_edges, n_branches
 * receiving their count. */
int lsc_network_advanced(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                         const double* source, const double* ground, int policy, const lsc_solve_options* options,
                         double* voltage, double* current, int32_t* branch_r, int32_t* branch_c,
                         double* branch_current, int64_t* n_branches, int32_t* solved, int32_t* n_components,
                         lsc_solve_stats* stats);

/* Receives a network pair's tables during a network pairwise run: the
 * component's nodes, their voltages relative to point ci (NULL unless
 * voltages were asked for) and node currents, and the branch currents. */
typedef void (*lsc_network_emit_fn)(int ci, int cj, int64_t n_nodes, const int32_t* nodes, const double* voltage,
                                    const double* current, int64_t n_branches, const int32_t* branch_r,
                                    const int32_t* branch_c, const double* branch_current, void* ctx);

/* Pairwise mode on a network (graph as in lsc_network_advanced): effective
 * resistances between the focal nodes `points` (-1: not in the graph),
 * excluded pairs (n_points x n_points uint8), every pair solved (no
 * shortcut). flags: 1 voltages, 2 currents, 4 cumulative currents, 8 max;
 * with any of them every pair's tables go to `emit`, and cum (n_nodes) and
 * the cumulative branch currents (capacity n_edges, count in
 * n_cum_branches) receive their sums over the pairs. Replaces
 * network_pairwise (src/cpp/circuit/python.cpp:348-402 -> raster_pairwise
 * with network set, circuitscape.cpp:474-743). resistances: n_points^2. */
int lsc_network_pairwise(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                         const int32_t* points, int64_t n_points, const uint8_t* excluded, int flags,
                         lsc_network_emit_fn emit, void* ctx, const lsc_solve_options* options, double* resistances,
                         double* cum, int32_t* cum_branch_r, int32_t* cum_branch_c, double* cum_branch_current,
                         int64_t* n_cum_branches, int32_t* n_components, lsc_solve_stats* stats);

/* Several advanced problems on one raster graph (the one-to-all driver).
 * Problem p: source cells scells[sptr[p]..sptr[p+1]) with values svals,
 * ground cells likewise, check cell check[p] (-1: every component).
 * Replaces circuitscape_multi_advanced (src/cpp/circuit/python.cpp:404-482 ->
 * raster_advanced_many, circuitscape.cpp:904-1083). Outputs per problem and
 * cell (nprob x ny x nx): voltage, current, active (uint8); solved[p];
 * nodemap (ny x nx). */
int lsc_circuitscape_multi_advanced(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                                    int avg_resistances, int four_neighbors, int policy, int64_t nprob,
                                    const int64_t* sptr, const int64_t* scells, const double* svals,
                                    const int64_t* gptr, const int64_t* gcells, const double* gvals,
                                    const int64_t* check, const lsc_solve_options* options, int32_t* nodemap,
                                    double* voltage, double* current, uint8_t* active, int32_t* solved,
                                    lsc_solve_stats* stats);

/* The same problems reduced on the device to what run_onetoall keeps when
 * no per-point map is asked for (src/landschaft/circuit/circuitscape.py:
 * 826-848): every problem's node currents (0 at its sources and grounds with
 * focal_zero) summed into cum and maximised into max per cell, in problem
 * order. cum (ny x nx) and max (ny x nx or NULL) are read and updated in
 * place, so calls over several polygon maps continue one sum. Per problem:
 * the voltage at the check cell's node (0 if not solved there) and solved[p].
 * Replaces, for that case, the per-problem outputs of
 * circuitscape_multi_advanced (src/cpp/circuit/python.cpp:404-482); one
 * problem's maps never leave the GPU. */
int lsc_circuitscape_multi_advanced_sum(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                                        int avg_resistances, int four_neighbors, int policy, int64_t nprob,
                                        const int64_t* sptr, const int64_t* scells, const double* svals,
                                        const int64_t* gptr, const int64_t* gcells, const double* gvals,
                                        const int64_t* check, int focal_zero, const lsc_solve_options* options,
                                        int32_t* nodemap, double* cum, double* max, double* check_voltage,
                                        int32_t* solved, lsc_solve_stats* stats);

/* J = sum_k R_eff(src_k, gnd_k) on the grid of conductance 1 / r (cells with
 * a finite r > 0, average conductance, 8 or 4 neighbours) and dJ/dr for every
 * cell by the adjoint (lambda_k = v_k). No reference counterpart (new:
 * resistance-surface calibration). R receives the npairs resistances, grad
 * ny x nx. */
int lsc_effective_resistance_grad(const double* resistance, int64_t nx, int64_t ny, int four_neighbors,
                                  const int64_t* src_cells, const int64_t* gnd_cells, int64_t npairs,
                                  const lsc_solve_options* options, double* J, double* R, double* grad,
                                  lsc_solve_stats* stats);

/* A factorised grid circuit kept for autograd (no reference counterpart):
 * create factors the grid of conductance 1 / r (host resistance), every
 * component grounded at its last node; solve maps k right-hand sides given
 * per cell (device, k x ny x nx) to potentials per cell (device); adjoint
 * solves L lambda = dJ/dV (device, k x ny x nx) with the same factor and
 * writes dJ/dr (device, ny x nx) from the potentials V. */
typedef struct lsc_grid_circuit lsc_grid_circuit;
int lsc_grid_circuit_create(const double* resistance, int64_t nx, int64_t ny, int four_neighbors,
                            const lsc_solve_options* options, lsc_grid_circuit** out);
/* update factorises the circuit again for a new host resistance on the same
 * cells (finite r > 0 where they were, status 1 otherwise), keeping the
 * symbolic plans: a calibration loop pays the analysis once. */
int lsc_grid_circuit_update(lsc_grid_circuit* circuit, const double* resistance);
int lsc_grid_circuit_solve(lsc_grid_circuit* circuit, int k, const double* rhs_cells, double* out_cells);
int lsc_grid_circuit_adjoint(lsc_grid_circuit* circuit, int k, const double* dJdV_cells, const double* V_cells,
                             double* grad_r);
int lsc_grid_circuit_stats(lsc_grid_circuit* circuit, lsc_solve_stats* stats);
void lsc_grid_circuit_destroy(lsc_grid_circuit* circuit);

/* The body of an ESRI ASCII grid of ny x nx float64 values (row-major):
 * every value in the shortest text that reads back to the same double (as
 * C++17 std::to_chars writes it), NaN written as -9999, one space between
 * values and a newline after each row. Replaces ascii_grid_rows
 * (src/cpp/circuit/python.cpp:484-521, registered at :554; the writer of
 * src/landschaft/circuit/circuitscape.py:1027 calls it). With out NULL,
 * *size receives the number of bytes of the text; otherwise the text goes to
 * out (capacity bytes, no terminating zero) and *size receives its length,
 * status 1 when capacity is too small. Host code: no device is needed. */
int lsc_ascii_grid_rows(const double* values, int64_t nx, int64_t ny, char* out, int64_t capacity, int64_t* size);

/* Symbolic statistics of a window plan with the reference's ordering
 * (16-cell nested-dissection leaves; disk: 1 for the circuit windows, 0 for
 * the flow-potential windows): out[7] = {unknowns, fronts, packed L entries,
 * factor flops (sum np^3/3 + np^2 nu + np nu^2), largest front rows, levels,
 * L entries as m x np blocks}. No reference counterpart (the reference's
 * figures of the same quantities are internal to frontal_analyze). */
int lsc_window_plan_stats(int radius, int buffer, int four_neighbors, int disk, double* out);

/* The symbolic analysis of a circuit: its node graph and the multifrontal
 * plan of every connected component's Laplacian, as the solves build them.
 * Replaces the analysis half of the reference's factor_component
 * (src/cpp/circuit/circuitscape.cpp:92-141: the graph from raster_graph,
 * :145-263, or network_graph, :265-338, then frontal_analyze_graph,
 * src/cpp/circuit/system.cpp:373-440), exposed on its own so that a caller
 * can size a job (nodes, factor entries, largest front) before any device
 * work. Host code: no device is needed. The digests are 64-bit hashes of the
 * node map, adjacency, weights and component labels (graph_digest) and of
 * every plan's elimination order, fronts, update rows, child-to-parent row
 * maps and assembled entries (plan_digest); equal inputs give equal digests
 * whatever the host thread count (LSC_HOST_THREADS, 1 to 16). */
typedef struct lsc_analysis {
    int64_t n_nodes;        /* graph nodes (short-circuit regions merged) */
    int64_t n_components;
    int64_t n_entries;      /* adjacency entries, both directions */
    int64_t n_fronts;       /* fronts of all components' plans */
    int64_t factor_entries; /* stored entries of L, blocks padded to 4 */
    int64_t largest_front;  /* rows of the largest front */
    double factor_flops;    /* sum np^3/3 + np^2 nu + np nu^2 over fronts */
    double graph_ms;        /* host wall time of the graph and its components */
    double plan_ms;         /* host wall time of the component matrices and plans */
    uint64_t graph_digest;
    uint64_t plan_digest;
} lsc_analysis;

/* Raster circuit: the arguments of lsc_circuitscape_advanced's graph. */
int lsc_circuitscape_analyze(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                             int avg_resistances, int four_neighbors, lsc_analysis* out);
/* Network circuit: the arguments of lsc_network_advanced's graph. */
int lsc_network_analyze(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                        lsc_analysis* out);

/* Checks the MFMA operand layout the front kernels assume: out[256] receives
 * the 16 x 16 product A (16x4) * B (4x16) of a = [i + 16 k], b = [k + 4 j]
 * computed with v_mfma_f32_16x16x4f32 (precision 0) or the f64 MFMA (1),
 * row-major. No reference counterpart (a self-test). */
int lsc_probe_mfma(int precision, double* out);

#ifdef __cplusplus
}
#endif

#endif /* LANDSCHAFT_CIRCUIT_H */
