// clb_internal.h -- shared declarations of the HIP path (context, kernel launchers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/clb.h"

namespace clb {

// ------------------------------------------------------------------ lattice tables
// D3Q19 velocity set; opposite pairs are (1,2), (3,4), ... (17,18).
#define CLB_D3Q19_CX {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0}
#define CLB_D3Q19_CY {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1}
#define CLB_D3Q19_CZ {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1}
// D3Q7: rest, +x, -x, +y, -y, +z, -z; weights 1/4 and 1/8 (c_s^2 = 1/4).
#define CLB_D3Q7_CX {0, 1, -1, 0, 0, 0, 0}
#define CLB_D3Q7_CY {0, 0, 0, 1, -1, 0, 0}
#define CLB_D3Q7_CZ {0, 0, 0, 0, 0, 1, -1}

__host__ __device__ constexpr int opp_dir(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

// ------------------------------------------------------------------ kernel argument blocks
struct Geom {
  int nx, ny, nz;
  long long N;
};

template <typename R>
struct FlowArgs {
  // 19*N, population-major. Slot (i, m) holds f_i - t_i (1 + rref(x_m)): the deviation from
  // the weights at the reference density 1 + rref(x) of the slot's own voxel, where
  // rref(x) = rref0 + x * rdx is the linear profile of the flow's initial state.
  R* f;
  const uint8_t* mask;    // N
  const uint8_t* kind;    // 256
  const R* omega;         // 256
  Geom g;
  R rhobar_in, rhobar_out;  // the faces' densities - 1
  R dface_in, dface_out;    // the same, less rref at x = 0 and x = nx-1
  R rdx;                    // rref(x + 1) - rref(x)
  double rref0, rrdx;       // rref(0) and rdx in FP64, for the reported densities
  // 1: no-dynamics voxels (kind NONE) of 1 <= x <= nx-2 stream their populations unchanged
  // (Palabos' NoDynamics: no collision, the lattice's streaming still moves them), because a
  // fluid voxel neighbours one; 0: no fluid voxel does, and they are skipped
  int stream_none;
  double* energy;         // [0..1] sum |u|^2/2 and count; [2 + 2b..] per-block partials; or null
  // the compact layout (clb_flow_compact.hip), when f holds its populations: node j's voxel
  // fnode[j], its slot at +c_d fnbr[(d-1) fcap + j] (d = 1..18), the node count *fnf, plane
  // stride fstride = fcap + ghost capacity; unused by the dense kernels
  const int* fnode;
  const int* fnbr;
  const unsigned int* fnf;
  long long fcap, fstride;
  // when set and *stop != 0 the step does nothing (the device tracer has converged)
  const int* stop;
};

// The device ValueTracer of clb_flow_resolve; its window (w doubles) is a separate array.
struct FlowTracer {
  double eps;
  int w;            // window length, (int)(L / u / 2)
  int n;            // values held
  int head;         // slot of the oldest
  int stop;         // 1 once converged: the steps launched after it do nothing
  long long steps;  // energies taken since clb_flow_tracer
  long long call0;  // steps at the start of the current resolve
};

// The pore <-> biofilm rule of updateLocalMaskNtotalLattices3D
// (src/complab3d_processors_part2.hh:99-210): the pore types (pore[0] the type a biofilm voxel
// returns to by default), each biomass field's biofilm code, and whether the biofilm label is
// the sum of the codes of the fields present (the reference's rule) or bio[0] (one biofilm
// code).
struct MaskRule {
  unsigned long long pore_bits[4];  // bit m: code m is a pore type
  int npore;
  int pore[CLB_MAX_PORE_CODES];
  int bio[CLB_MAX_BIOMASS];
  int sum_labels;
};

__host__ __device__ inline bool mr_is_pore(const MaskRule& r, int m) {
  return ((r.pore_bits[(m >> 6) & 3] >> (m & 63)) & 1ull) != 0;
}

template <typename R>
struct AdeArgs {
  R* g;                   // S*7*N population-major; stored as f_i - t_i*c_ref[s]
  uint8_t* mask;          // updated in place by the biomass threshold rule
  const uint8_t* kind;    // 256: the solute lattices' dynamics per mask code (soluteDomainSetup)
  const uint8_t* flow_kind;  // 256: the flow lattice's (NSdomainSetup); node conversion only
  const R* omega;         // S*256
  const R* u;             // 3*N (x, y, z planes)
  R* bmass;               // nb*N
  double* cbuf;           // S*N post-kinetics concentrations for the speciation kernel (or null)
  // Speciation increments deferred to the next load (fused speciation only, else null):
  // pend_r holds the previous step's dC per solute and voxel, pend_w receives this step's.
  const R* pend_r;
  R* pend_w;
  // Mineral dissolution (else null): concentrations after the biotic kinetics of every
  // reactive voxel (the water a neighbouring mineral sees), S*N.
  R* water;
  // Surrogate metabolic model (else null): the processor's increments of this step,
  // S planes of dC then nb planes of dB, written before transport for every reactive voxel
  // holding a surrogate organism (a biomass field of srg_pools above 1e-12), read only there.
  const R* srg_inc;
  int srg_pools;          // bit b: biomass field b is grown by the surrogate
  Geom geo;
  // Compact (fluid-only) layout, when cnode is set (clb_compact.hip). Then g, u, pend_r,
  // pend_w and cbuf hold the ADE-fluid voxels only, in natural order (j = 0 .. nf-1), with
  // stride ccap (populations: cstride), while mask and biomass stay dense (index cnode[j]).
  // A population array (s, i) has ccap node slots, then the ghost slots of direction i: for
  // a fluid node whose neighbour at +c_i is a wall, the wall's slot i of the dense layout
  // (the bounce-back link's storage).
  const int* cnode;          // [ccap] dense index of fluid node j
  const int* cnbr;           // [6 * ccap] direction d = 1..6 at (d-1)*ccap + j: the slot at +c_d,
                             //   a fluid node's j (< ccap), a ghost slot (>= ccap), or -1 beyond an x face
  const unsigned int* cnf;   // number of fluid nodes (device memory, so captured steps survive a rebuild)
  long long ccap, cstride;
  int S, nb;
  int bc_type[2 * CLB_MAX_SOLUTES];
  R bc_value[2 * CLB_MAX_SOLUTES];
  R c_ref[CLB_MAX_SOLUTES];
  int immobile[CLB_MAX_SOLUTES];  // 1: not transported, whole concentration in the rest slot
  int kinetics;           // 0 none, 1 plugin
  int abiotic;            // 0 none, 1 plugin everywhere, 2 plugin at pore voxels next to a wall
  double dt;
  int solid_code, bb_code;
  int stream_none;        // as FlowArgs::stream_none, for the solute lattices' kinds
  int mask_update;        // the rule inside the transport kernel (one pore type only)
  double bthreshold;
  int pore_code, biofilm_code;
  MaskRule mrule;
};

struct Tableau {
  int nc, ns;
  double S[CLB_MAX_SPECIES * CLB_MAX_COMPONENTS];
  double logK[CLB_MAX_SPECIES];
  double inv_mu0[CLB_MAX_COMPONENTS];  // 1 / (smallest positive coefficient <= 1)
  int is_eq[CLB_MAX_SPECIES];          // row has a nonzero entry
  double eqw[CLB_MAX_SPECIES];         // 1.0 for an equilibrium species (i < ns), else 0.0
  int comp_species[CLB_MAX_COMPONENTS];
  double tol;
  int max_iter;
};

// Batched / coupled speciation problem description.
struct SpecArgs {
  long long n;               // number of solves (batch) or nodes (coupled)
  const double* totals;      // n*nc (batch) or null
  const double* logk_var;    // n*nvar or null
  int nvar;
  int var_species[4];
  double* logc;              // n*nc in/out
  int* iters;                // n
  double* resid;             // n
  uint8_t* conv;             // n
  unsigned int* counter;     // work-queue head
  unsigned long long* stats; // solves, converged, iterations
  const unsigned int* n_dev; // item count in device memory (overrides n when set)
  unsigned long long* blk_stats;  // fused step: per-wavefront statistics pairs, and their number
  long long blk_stats_n;
  int straggler_evals;            // fused step: budget of the stragglers' first pass (0: none)
  // fused step: the stragglers as the transport kernel left them, by segment (segment s holds
  // seg_n[s] of them from s * segcap on), and where the first pass leaves what it does not
  // solve for the full iteration (fail_n of them)
  const int* seg_list;
  const double* seg_x;
  const unsigned int* seg_n;
  long long segcap;
  int* fail_list;
  double* fail_x;
  unsigned int* fail_n;
};

// The fused step's stragglers are appended to kStragglerSegments segments of the list, block b
// of the transport kernel to segment b mod kStragglerSegments, each with a counter of its own:
// atomics on one address serialise across the device, and one counter for every wavefront with
// a straggler made the appends a bottleneck of the transport kernel. A segment holds at most
// the voxels of its blocks: segcap = ceil(blocks / kStragglerSegments) x threads per block.
constexpr int kStragglerSegments = 64;

// Transport step with the speciation fused in (clb_ade_spec.h).
struct FusedSpecIO {
  int* slist;                     // voxels left to the persistent kernel, by segment
  unsigned int* scount;           // kStragglerSegments counters (cleared before the step)
  unsigned long long* blk_stats;  // per wavefront of the launch: solves, evaluations
  double* sx;                     // per straggler: log activities its solve continues from (stride 4)
  int evals;                      // evaluation budget of the fused solve (>= 2)
  long long segcap;               // capacity of a segment
};

// ------------------------------------------------------------------ launchers
template <typename R>
void launch_flow_init(const Geom& g, R* f, hipStream_t s);
// flag[0] = 1 when a voxel of kind NONE has a voxel of kind FLUID among its neighbours in
// the D3Q19 (q = 19) or D3Q7 (q = 7) stencil; flag[0] is left alone otherwise.
void launch_none_reach(const uint8_t* mask, const uint8_t* kind, const Geom& g, int q, unsigned int* flag,
                       hipStream_t s);
// Per-block energy partials a launch of this block shape writes (the energy buffer holds
// 2 + 2 * flow_blocks doubles).
long long flow_blocks(const Geom& g, dim3 block);
// energy[0..1] = the sums of the nblocks per-block partials, in a fixed order
void launch_energy_reduce(double* energy, long long nblocks, hipStream_t s);
// the device tracer after a flow step (energy: that step's partials; elog: the energies of the
// current resolve, or null) and at the start of a resolve
void launch_tracer_step(const double* energy, long long nblocks, FlowTracer* tr, double* win, double* elog,
                        hipStream_t s);
void launch_tracer_begin(FlowTracer* tr, hipStream_t s);
// with a.energy: the per-block partials, summed into energy[0..1] when reduce (the device tracer
// sums them itself)
template <typename R>
void launch_flow_step(const FlowArgs<R>& a, bool odd, dim3 block, hipStream_t s, bool reduce = true);
template <typename R>
void launch_flow_velocity(const FlowArgs<R>& a, bool odd_next, double* u, double* rho, R* ade_u, hipStream_t s);

template <typename R>
void launch_ade_init(const Geom& g, int S, R* gpop, const double* c0, const R* c_ref, const int* immobile,
                     hipStream_t s);
template <typename R>
void launch_mask_rule(const R* b, int nb, const uint8_t* prev, uint8_t* mask, const uint8_t* kind, const Geom& g,
                      int solid, int bb, const MaskRule& r, double thr, hipStream_t s);
// changed[0] += voxels whose code changed; changed[1] += those of them that were walls for the
// flow and are fluid now (their flow populations written into flowf, the dense array)
template <typename R>
void launch_node_conversion(const AdeArgs<R>& a, bool odd_next, uint8_t* phase, int ms, int pore_code,
                            int precip_code, int phase_id, double precip_max, double reopen, unsigned int* changed,
                            R* flowf, bool flow_next_odd, double fref0, double frdx, hipStream_t s);
template <typename R>
void launch_ade_reset(const AdeArgs<R>& a, bool next_odd, const double* c0, hipStream_t s);
template <typename R>
void launch_solute_field(const AdeArgs<R>& a, bool next_odd, int solute, const double* c, hipStream_t s);
template <typename R>
void launch_flow_stats(const FlowArgs<R>& a, bool odd_next, double* acc, hipStream_t s);
template <typename R>
void launch_ade_step(const AdeArgs<R>& a, bool odd, dim3 block, hipStream_t s);
template <typename R>
void launch_ade_conc(const AdeArgs<R>& a, bool odd_next, double* c, hipStream_t s);
// Speciation on the coupled step: solve at every reactive node from cbuf and add the
// limited increments to the populations the step just wrote.
void launch_spec_coupled(const AdeArgs<float>& a, bool odd_done, const Tableau& tb, SpecArgs sa, int grid,
                         hipStream_t s);
void launch_spec_coupled(const AdeArgs<double>& a, bool odd_done, const Tableau& tb, SpecArgs sa, int grid,
                         hipStream_t s);

// Compact layout (clb_compact.hip). Build: fluid flags and bounce-back links of every voxel,
// scans, node table. The caller sizes ccap / ghost capacity from the counts.
// The order of the compact nodes: bricks of bx x by x bz voxels in x-fastest order, natural
// inside each (bx = nx, by = bz = 1: the natural order). nkey = nbx nby nbz vol >= N.
struct BrickOrder {
  int bx, by, bz, nbx, nby, nbz;
  long long vol, nkey;
};
struct CompactBuild {
  unsigned int* flag;        // [nkey] 1: fluid (brick order)
  unsigned int* cj;          // [nkey] exclusive scan of flag
  unsigned int* lk;          // [6 nkey + 1] 1: fluid voxel with a wall at +c_d (direction-major)
  unsigned int* cg;          // [6 nkey + 1] exclusive scan of lk
  void* scan_tmp;            // scan workspace
  size_t scan_tmp_bytes;
};
size_t compact_scan_bytes(long long n);  // workspace the scans need for n items
void launch_compact_flags(const uint8_t* mask, const uint8_t* kind, const Geom& g, const BrickOrder& o,
                          const CompactBuild& b, hipStream_t s);
void launch_compact_scan(const CompactBuild& b, long long nkey, hipStream_t s);
// node table and the fluid count (written to cnf)
void launch_compact_table(const CompactBuild& b, const Geom& g, const BrickOrder& o, int* cnode, int* cnbr,
                          long long ccap, unsigned int* cnf, hipStream_t s);
// Dense <-> compact copies of the solute state, for the fluid nodes of ca's table:
// populations with the ghost slots (to_compact: from g_dense into ca.g; else back), both
// deferred-increment buffers (2 * S planes each, dense stride N, compact stride ccap; null:
// none) and, towards the compact layout, the velocity (u_dense into ca.u).
template <typename R>
void launch_compact_move(bool to_compact, const AdeArgs<R>& ca, R* g_dense, const R* u_dense, R* pend_dense,
                         R* pend_compact, hipStream_t s);
// compact -> dense with the deferred increments (ca.pend_r / pend_w, the compact buffers)
// folded in, for a next step of even parity; the dense increments come out zero
template <typename R>
void launch_compact_move_flush(const AdeArgs<R>& ca, R* g_dense, R* pend_dense, R* pend_compact, hipStream_t s);
// pending_flush on the compact layout (a.pend_r / pend_w: the compact buffers, as for the
// next step of parity odd_next)
template <typename R>
void launch_compact_flush(const AdeArgs<R>& a, bool odd_next, hipStream_t s);
template <typename R>
void launch_ade_step_compact(const AdeArgs<R>& a, bool odd, int bt, hipStream_t s);

// Compact flow layout (clb_flow_compact.hip): the voxels of flow kind FLUID in brick order,
// with per direction (d = 1..18) the slot at +c_d: a flow node's index, a ghost slot (>= fcap,
// plane d) where the neighbour is a wall, -1 beyond an x face. The build buffers are
// CompactBuild's with lk / cg of 18 nkey + 1 entries.
size_t flow_compact_scan_bytes(long long nkey);
void launch_flow_compact_build(const uint8_t* mask, const uint8_t* kind, const Geom& g, const BrickOrder& o,
                               const CompactBuild& b, hipStream_t s);
void launch_flow_compact_table(const CompactBuild& b, const Geom& g, const BrickOrder& o, int* fnode, int* fnbr,
                               long long fcap, unsigned int* fnf, hipStream_t s);
// dense <-> compact copies of the flow populations (a.f: the compact array; f_dense: 19 N):
// every node's own slots and the wall slots behind its links
template <typename R>
void launch_flow_compact_move(bool to_compact, const FlowArgs<R>& a, R* f_dense, hipStream_t s);
// per-block energy partials of a compact launch of bt threads
long long flow_compact_blocks(long long fcap, int bt);
template <typename R>
void launch_flow_step_compact(const FlowArgs<R>& a, bool odd, int bt, hipStream_t s, bool reduce = true);
// the advecting velocity of the solutes (latticeToPassiveAdvDiff) from the compact flow
// state: every node of the table writes ade_u at its voxel (the caller zeroes the rest)
template <typename R>
void launch_flow_velocity_compact(const FlowArgs<R>& a, bool odd_next, R* ade_u, hipStream_t s);
template <typename R>
bool launch_ade_spec_step_compact(const AdeArgs<R>& a, bool odd, int bt, const Tableau& tb, const FusedSpecIO& io,
                                  hipStream_t s);
long long fused_waves_compact(long long ccap, int bt);

// Returns false when no fused shape is compiled for (S, nc): the caller then uses the
// transport kernel + spec_kernel over the reactive list.
template <typename R>
bool launch_ade_spec_step(const AdeArgs<R>& a, bool odd, dim3 block, const Tableau& tb, const FusedSpecIO& io,
                          hipStream_t s);
// Wavefront slots of a fused launch (the per-wavefront statistics it writes): blocks of at
// most 256 threads hold 4, larger blocks (at most 512 threads) 8.
long long fused_waves(const Geom& g, dim3 block);
// Add the deferred speciation increments (pend_r of a) to the stored populations, laid out
// for a next step of parity odd_next.
template <typename R>
void launch_pending_flush(const AdeArgs<R>& a, bool odd_next, hipStream_t s);

// grid: spec_kernel's persistent grid; grid8: spec_group_kernel's (tableaux of more than
// four components)
void launch_spec_batch(const Tableau& tb, SpecArgs sa, int warm, int method, int grid, int grid8, hipStream_t s);
int spec_occupancy_grid(int nc, int device);
int spec_group_grid(int device);

// Mineral dissolution (clb_dissol.hip): surfaceDissolutionKinetics3D with the compiled
// defineDissolutionRate hook. phase_sub[id] = the immobile solute holding phase id's
// inventory (-1: not dissolvable), ids 1..CLB_MAX_PHASES-1.
#define CLB_MAX_PHASES 16
struct DissolArgs {
  const uint8_t* phase;       // N phase ids (0: none)
  int phase_sub[CLB_MAX_PHASES];
  int surface_only;           // 1: a phase voxel without an open face neighbour does not react
};
template <typename R>
void launch_dissolution(const AdeArgs<R>& a, const DissolArgs& d, bool odd_done, hipStream_t s);

// Surrogate metabolic model (clb_surrogate.hip): run_surrogate3D with a MATLAB-exported
// feed-forward network per microbe. Tables live in device memory (uniform reads).
#define CLB_SRG_MAX_WIDTH 16
#define CLB_SRG_MAX_LAYERS 6
struct SrgNet {
  int on;                       // 0: the microbe has no network (grows at 0, decays)
  int n_in, n_layers, log_output, closure;
  int width[CLB_SRG_MAX_LAYERS];
  int in_sub[CLB_MAX_SOLUTES];  // input k reads Fin[microbe][in_sub[k]]
  double x_offset[CLB_MAX_SOLUTES], x_gain[CLB_MAX_SOLUTES];
  double x_ymin, y_ymin, y_gain, y_offset;
  double yield[CLB_MAX_SOLUTES];  // gDW per mmol (closure 1, 2); <= 0: not tied to growth
  double W[CLB_SRG_MAX_LAYERS][CLB_SRG_MAX_WIDTH][CLB_SRG_MAX_WIDTH];  // row = neuron, zero padded
  double b[CLB_SRG_MAX_LAYERS][CLB_SRG_MAX_WIDTH];
};
struct SrgTables {
  int nm;                                   // surrogate microbes
  int id[CLB_MAX_BIOMASS];                  // their biomass fields
  double Kc[CLB_MAX_BIOMASS][CLB_MAX_SOLUTES];
  double vmax[CLB_MAX_BIOMASS][CLB_MAX_SOLUTES];   // 0: unset
  double maxup[CLB_MAX_BIOMASS][CLB_MAX_SOLUTES];  // floor (<= 0); <= -1e29: unset
  double decay[CLB_MAX_BIOMASS];            // 1/s
  double mw[CLB_MAX_BIOMASS];               // gDW per mol
  int fixc[CLB_MAX_SOLUTES], fixlb[CLB_MAX_SOLUTES];
  int use_totals;                           // uptake limited by component totals
  int totals_active;                        // the tableau maps at least one component
  int has_contrib[CLB_MAX_SOLUTES];
  double contrib[CLB_MAX_SOLUTES][CLB_MAX_SOLUTES];  // [substrate][species] coefficient, 0: none
  SrgNet net[CLB_MAX_BIOMASS];
  int lds_layers;                           // layers per network in the kernel's LDS copy (max n_layers)
};
// Evaluates the processor at every reactive voxel holding a surrogate organism, writing
// its S + nb increments to inc.
// nm, lds_layers: the tables' values (they size the kernel's LDS copy of the networks).
template <typename R>
void launch_surrogate(const AdeArgs<R>& a, const SrgTables* t, int nm, int lds_layers, R* inc, bool odd,
                      hipStream_t s);

// CA biomass redistribution (clb_ca.hip)
void launch_ca_init(const uint8_t* mask, int* age, uint8_t* dpos, const Geom& g, int solid, int bb, const MaskRule& r,
                    hipStream_t s);
template <typename R>
void launch_ca_mask_total(const R* b, int nb, const uint8_t* prev, uint8_t* mask, double* btot, const Geom& g,
                          int solid, int bb, const MaskRule& r, double bthr, bool fresh, unsigned long long* red,
                          hipStream_t s);
void launch_ca_push(const uint8_t* mask, const double* btot, const int* age, double* out, uint8_t* flags,
                    const Geom& g, int solid, int bb, const MaskRule& r, double bmax, int half,
                    unsigned long long seed, unsigned long long sweep, hipStream_t s);
template <typename R>
void launch_ca_gather(const R* b_old, R* b, int nb, const uint8_t* mask, const double* btot, const double* out,
                      const uint8_t* flags, const Geom& g, int solid, int bb, hipStream_t s);
void launch_ca_age(const int* age_in, int* age_out, const uint8_t* dpos, const double* btot, const Geom& g,
                   double bmax, unsigned long long* changed, hipStream_t s);
// The FD biofilm solver's explicit diffusion step (fdDiffusion3D) from the copy b_old into b.
template <typename R>
void launch_fd_biomass(const R* b_old, R* b, int nb, const uint8_t* mask, const Geom& g, double nu, hipStream_t s);

// MLP kernels (FP32 training, BF16 MFMA inference)
struct MlpDims {
  int din, H, dout;
};
// Fused training step (forward, loss gradient, backward) for H = 128, din <= 16, dout <= 16:
// one workgroup per mlp_step_rows() rows of X / T from row0 on (B rows), each writing its
// partial gradient to grad + z gs (gs >= P and a multiple of 4).
bool mlp_step_fits(const MlpDims& d);
int mlp_step_rows();
void launch_mlp_step(const MlpDims& d, long long B, long long row0, const float* X, const float* T,
                     const float* theta, float* grad, long long gs, float scale, double* loss, hipStream_t s);
// out[i] (+)= sum_z g[z gs + i] for i < P, z ascending
void launch_partial_sum(long long P, const float* g, int splits, long long gs, float* out, int accumulate,
                        hipStream_t s);
void launch_gemm(int transA, int transB, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                 float* C, int ldc, const float* bias, int act, float beta, hipStream_t s);
// K split in `splits` chunks, chunk z written to C + z sc
void launch_gemm_split(int transA, int transB, int M, int N, int K, const float* A, int lda, const float* B,
                       int ldb, float* C, int ldc, const float* bias, int act, float beta, int splits, long long sc,
                       hipStream_t s);
void launch_tanh_backward(long long n, float* dz, const float* h, hipStream_t s);
// column sums over rows split in `splits` chunks, chunk z written to out + z so
void launch_colsum(int rows, int cols, const float* A, float* out, int splits, long long so, hipStream_t s);
void launch_mse_grad(long long n, const float* Y, const float* T, float scale, float* dY, double* loss,
                     hipStream_t s);
// gradient = sum of `splits` partials at g + z gs
void launch_adam(long long n, float* p, const float* g, int splits, long long gs, float* m, float* v, float lr,
                 float b1, float b2, float eps, float b1t, float b2t, hipStream_t s);
void launch_mlp_infer_bf16(const MlpDims& d, long long n, const float* theta, const float* X, float* Y,
                           hipStream_t s);
void launch_mlp_warm(const MlpDims& d, long long n, const float* theta, const double* xmin, const double* xmax,
                     const double* ymin, const double* ymax, const Tableau& tb, const SpecArgs& sa, hipStream_t s);
void launch_mlp_make_dataset(const MlpDims& d, long long n, const Tableau& tb, const SpecArgs& sa, float* X,
                             float* T, const double* xmin, const double* xmax, const double* ymin,
                             const double* ymax, hipStream_t s);
void launch_fill_u32(unsigned int* p, long long n, unsigned int v, hipStream_t s);
void launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s);
void launch_fill_f64(double* p, long long n, double v, hipStream_t s);

}  // namespace clb

#define CLB_HIP_LAUNCH_CHECK() ((void)0)
