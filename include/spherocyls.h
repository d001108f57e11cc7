/*
 * spherocyls.h -- C ABI of libspherocyls.so, the MI355X (gfx950) spherocylinder
 * dynamics library.
 *
 * The reference (tmarschall/spherocyls) has no FFI: its boundary is the C++ class
 * SpherocylBox (src/cpp/spherocyl_box.h:85-120), whose methods the entry points
 * below replace one by one (each cites the method it stands for).  The C++ host
 * class in include/spherocyl_box.h keeps that class's names and
 * semantics on top of this ABI.
 *
 * Conventions
 *   - Every entry point returns an int status (SC_OK = 0); nothing calls exit().
 *   - All arrays are caller-owned host memory, copied in and out; particle arrays
 *     are in the caller's particle order (the library sorts particles by cell on
 *     the device and undoes the permutation on every copy-out).
 *   - A system holds one or more independent packings ("a batch"); packing k owns
 *     the particles offsets[k] .. offsets[k+1]-1 of the concatenated arrays.  The
 *     reference's single box is a system with one packing.
 *   - One system per host thread; calls on one system are serialised on its own
 *     HIP stream.  No global mutable state is shared between systems.
 */
#ifndef SPHEROCYLS_H
#define SPHEROCYLS_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SC_OK 0
#define SC_ERR_ARG 1            /* invalid argument */
#define SC_ERR_HIP 2            /* HIP runtime failure (message in sc_last_error) */
#define SC_ERR_NO_DEVICE 3      /* no usable GPU */
#define SC_ERR_CELL_OVERFLOW 4  /* more particles in one cell than the bucket holds */
#define SC_ERR_PLACEMENT 6      /* random placement gave up */
#define SC_ERR_LIST_OVERFLOW 7  /* more than the per-particle candidate list holds */
#define SC_ERR_NOT_CONVERGED 8  /* minimiser hit its iteration limit */
#define SC_ERR_STATE 9          /* call not valid in the current state */
#define SC_ERR_HOST 10          /* host-side failure: out of host memory or an internal
                                   error (message in sc_last_error) */

/* precision modes (sc_set_precision) */
#define SC_PREC_FP64 0  /* all pair math in FP64 */
#define SC_PREC_MIXED 1 /* FP32 pair math on FP64 relative displacements for the
                           dynamics and the minimiser; FP64 sums and residual;
                           passes that record stress and energy (sc_calc_se,
                           se_every records) evaluate their pairs in FP64 */

/* sc_get_box fields (double out[SC_BOX_NFIELDS]) */
#define SC_BOX_LX 0
#define SC_BOX_LY 1
#define SC_BOX_GAMMA 2
#define SC_BOX_GAMMA_TOTAL 3
#define SC_BOX_E 4     /* energy, sum over contacts of (1-d/sigma)^2 / (2N) */
#define SC_BOX_PXX 5   /* stress with the reference's own formulas            */
#define SC_BOX_PYY 6   /*   (src/cpp/spherocyl_box.cpp:483-495)                */
#define SC_BOX_PXY 7
#define SC_BOX_CPXX 8  /* corrected virial P_ab = sum F_b (CA_a - CB_a) / area */
#define SC_BOX_CPYY 9
#define SC_BOX_CPXY 10
#define SC_BOX_CPYX 11
#define SC_BOX_NCONTACTS 12
#define SC_BOX_FRMS 13 /* sqrt(sum(Fx^2+Fy^2+tau^2)/(3N)) of the last force pass */
#define SC_BOX_PACKING 14
#define SC_BOX_ROWS 15
#define SC_BOX_COLS 16
#define SC_BOX_PAD 17
#define SC_BOX_ITERS 18 /* minimiser iterations of the last sc_fire_minimize */
#define SC_BOX_FIRE_DT 19    /* FIRE time step and mixing factor at the end of the */
#define SC_BOX_FIRE_ALPHA 20 /*   last sc_fire_minimize                            */
#define SC_BOX_PAIR_FP32 21  /* 1 while this packing's pair math runs in FP32      */
#define SC_BOX_NFIELDS 22

/* sc_strain_steps / sc_resize_steps record layout: per recorded sub-step and
   packing, SC_SE_NFIELDS doubles: E Pxx Pyy Pxy (reference formulas),
   cPxx cPyy cPxy cPyx (corrected), contacts */
#define SC_SE_NFIELDS 9

typedef struct sc_system sc_system;

int sc_version(void);
const char* sc_status_string(int status);
/* Number of visible GPUs (0 on a host without one; never fails). */
int sc_device_count(void);

/* SpherocylBox(dL, N, pdX, pdY, pdPhi, pdA, pdR, dPadding)
 *   src/cpp/spherocyl_box.cpp:152-173 -- copies the arrays, wraps positions and
 *   angles as set_particle does (:188-221), sizes cells as configure_cells
 *   (:223-268) and bins particles (find_neighbors, :320-355).
 * Batched form: npack packings, offsets[npack+1], L[npack], pad[npack]. */
int sc_system_create(sc_system** out, int device, int npack, const int* offsets,
                     const double* L, const double* x, const double* y, const double* phi,
                     const double* A, const double* R, const double* pad);
/* ~SpherocylBox()  src/cpp/spherocyl_box.cpp:174-186 */
int sc_system_destroy(sc_system* s);
/* Text of the last error on this system ("" if none). */
const char* sc_last_error(const sc_system* s);
int sc_num_particles(const sc_system* s, int* n_total, int* npack);

/* set_strain_rate / set_resize_rate / set_step  src/cpp/spherocyl_box.h:112-114
   (k = -1 applies to every packing) */
int sc_set_dynamics(sc_system* s, int k, double strain_rate, double resize_rate, double step);
/* set_gamma / set_total_gamma  src/cpp/spherocyl_box.h:118-119 */
int sc_set_gamma(sc_system* s, int k, double gamma, double gamma_total);
int sc_set_precision(sc_system* s, int mode);
/* Contact model.  SC_CONTACT_CLOSEST: one contact at the closest points of the
 * two segments (the CPU reference, calc_force_pair src/cpp/strain_dynamics.cpp:
 * 17-75; the default).  SC_CONTACT_TWO_POINT: a second contact at the other end
 * of the overlap of nearly parallel rods, so aligned overlapping rods have no
 * energy kink (the reference's GPU two-point model, src/cuda/2point_dynamics.cu:
 * 391-546); needed for minimisation to F_rms < 1e-10 once rods pack side by
 * side.  Energy, stress and forces sum over the contact points. */
#define SC_CONTACT_CLOSEST 0
#define SC_CONTACT_TWO_POINT 1
int sc_set_contact_model(sc_system* s, int model);
/* Replay fixed launch sequences (the FIRE iteration chunks) as captured HIP
 * graphs (enable != 0, the default) or launch every kernel eagerly.  Results
 * are identical; graphs cut launch overhead for small systems.  Event timing
 * (sc_timing_enable) always launches eagerly. */
int sc_set_graphs(sc_system* s, int enable);
/* FIRE of small packings in one launch (no reference counterpart; a launch
 * choice like sc_set_graphs).  On the list path, once every packing that
 * sc_fire_minimize still advances holds at most max_particles particles
 * (0 <= max_particles <= 256; default 256, 0 = never), one workgroup per
 * packing runs the remaining iterations inside a single kernel: same
 * iteration, decisions and stop rule, sums in another order (results agree
 * to rounding, not bit for bit).  A launch runs at most steps_per_launch
 * iterations (0: the default, 50000) and the packings still running are
 * launched again, so no single kernel runs long.  With event timing on, that
 * kernel is counted under SC_KERNEL_FIRE_FORCE. */
int sc_set_small_fire(sc_system* s, int max_particles, int steps_per_launch);
int sc_get_contact_model(const sc_system* s, int* model);
/* Force-pass implementation (no reference counterpart; both give the same
 * forces to rounding):
 *   SC_PATH_LIST  one thread per particle over a per-particle partner list kept
 *                 with a skin (the cell padding) and rebuilt on the device when a
 *                 particle has moved pad/2 or the strain left the list's window;
 *   SC_PATH_TILE  one workgroup per run of cells, the 3 x (width+2) surrounding
 *                 cells staged in LDS, 8 lanes per particle with xor-shuffle sums. */
#define SC_PATH_TILE 0
#define SC_PATH_LIST 1
int sc_set_path(sc_system* s, int path);
int sc_get_path(const sc_system* s, int* path);
/* Tile path shape: columns per tile and threads per workgroup. */
int sc_set_tiling(sc_system* s, int tile_cols, int block_threads);
/* Active path's shape: tile width (tile path) or lanes per particle (list path),
   and its workgroup size. */
int sc_get_tiling(const sc_system* s, int* tile_cols, int* block_threads);
/* List path shape: workgroup size (64, 128 or 256) and lanes per particle
 * (1, 2, 4 or 8; more lanes give small systems more parallelism). */
int sc_set_list_block(sc_system* s, int block_threads, int lanes);
/* Cell padding (the list skin) of packing k (k = -1: all): the cell grid is
 * rebuilt for it (pad 0 takes the grid's slack, as at creation) and so is every
 * list row.  Results do not depend on the padding beyond the summation order
 * of the rows; the time per step does.  (The reference fixes the padding when
 * the box is built, SpherocylBox(..., dPadding), src/cpp/spherocyl_box.cpp:152-173.) */
int sc_set_pad(sc_system* s, int k, double pad);
/* Cell-size autotuner: for each of ncand candidate paddings, `steps` Heun steps
 * of the system's own dynamics from the current state (restored after every
 * trial), timed with their row rebuilds; keeps the fastest padding for every
 * packing.  ms_out[ncand] (may be NULL) gets each trial's milliseconds and
 * *chosen (may be NULL) the padding kept. */
int sc_autotune_pad(sc_system* s, int ncand, const double* pads, int steps, double* ms_out,
                    double* chosen);
/* Autotuner: times reps force passes of this system for each candidate path and
 * launch shape and keeps the fastest.  *shape gets the tile width (tile path)
 * or the lanes per particle (list path).  Outputs may be NULL. */
int sc_autotune(sc_system* s, int reps, int* path, int* shape, int* block_threads,
                double* us_per_pass);
/* Autotuner for minimisation workloads (no reference counterpart): for each
 * candidate path and launch shape, `iterations` FIRE iterations (params as in
 * sc_fire_minimize, max_iter replaced) from the current positions, restored
 * after every trial and at the end; keeps the fastest.  A minimiser's force
 * pass (two-point contacts, list rebuilds as particles move, the update
 * kernel) ranks the shapes differently from the Heun predictor that
 * sc_autotune times.  *ms gets the kept shape's wall time for the trial. */
int sc_autotune_fire(sc_system* s, const double* params, double switch_frms, int iterations,
                     int* path, int* shape, int* block_threads, double* ms);

/* set_particle for every particle (src/cpp/spherocyl_box.cpp:188-221): new
   positions / angles, wrapped, then cells rebuilt. */
int sc_set_positions(sc_system* s, const double* x, const double* y, const double* phi);
int sc_get_positions(sc_system* s, double* x, double* y, double* phi);
/* Per-particle force and torque of the last force pass (the reference keeps them
   in Spherocyl::m_dFx/m_dFy/m_dTau, src/cpp/spherocyl_box.h:17-19). */
int sc_get_forces(sc_system* s, double* fx, double* fy, double* tau);
int sc_get_box(sc_system* s, int k, double* out);

/* calc_forces  src/cpp/strain_dynamics.cpp:137-153 */
int sc_calc_forces(sc_system* s);
/* calc_se  src/cpp/spherocyl_box.cpp:502-522 (forces, energy, stress) */
int sc_calc_se(sc_system* s);

/* nsteps Heun shear steps: the inner loop of run_strain
 * (src/cpp/strain_dynamics.cpp:285-296, strain_step :179-235, Lees-Edwards remap
 * set_back_gamma src/cpp/spherocyl_box.cpp:524-545).  Energy and stress are
 * recorded after every se_every-th sub-step (0: never) into
 * se_out[rec][pack][SC_SE_NFIELDS]; forces are left at the final state. */
int sc_strain_steps(sc_system* s, int nsteps, int se_every, double* se_out);
/* nsteps Heun resize steps (resize_step, src/cpp/resize_dynamics.cpp:164-243):
 * shrink != 0 compresses.  corrected_image != 0 uses the rescaled box length in
 * the predictor pass (the reference uses the old one, :101). */
int sc_resize_steps(sc_system* s, int nsteps, int shrink, int corrected_image, int se_every,
                    double* se_out);
/* Pure shear, the GPU code's pure_strain_step_2p (src/cuda/2point_dynamics.cu:
 * 1663-2285) in this library's conventions (mobility 1, torque / I): each
 * step is a Heun step under the affine field that expands x and contracts y
 * at half the packing's strain rate.  Predictor: x~ = x + dt (Fx + rate x / 2),
 * y~ = y + dt (Fy - rate y / 2), phi~ = phi + dt (tau / I - rate RC sin(2 phi) / 2);
 * the box becomes Lx (1 + dt rate / 2) by Ly (1 - dt rate / 2); corrector:
 * x = x~ + dt (Fx~ - Fx) / 2 and likewise y and phi, the forces of the
 * predicted state evaluated in the new box.  gamma_total advances by
 * dt rate; the Lees-Edwards strain gamma stays.  List path only
 * (SC_ERR_STATE otherwise).  se_every / se_out as in sc_strain_steps. */
int sc_pure_strain_steps(sc_system* s, int nsteps, int se_every, double* se_out);
/* flip_shear_direction (src/cuda/find_neighbors.cu:552-579): x -> Lx - x,
 * phi -> -phi, gamma -> -gamma for packing k (k = -1: all). */
int sc_flip_shear(sc_system* s, int k);
/* rotate_by_gamma (src/cuda/find_neighbors.cu:580-645): a rigid rotation of the
 * lab frame that takes the sheared box vector (gamma Ly, Ly) to the new x axis.
 * Positions x' = sgn(gamma) y sqrt(1+gamma^2), y' = -sgn(gamma) x / sqrt(1+gamma^2)
 * (wrapped), phi' = phi - atan(1/gamma) (pi/2 at gamma = 0), box
 * Lx' = Ly sqrt(1+gamma^2), Ly' = Lx / sqrt(1+gamma^2), strain -gamma, for
 * packing k (k = -1: all).  Later shear steps remap at gamma = Lx/(2 Ly) with
 * x += (Lx/Ly) y (the general Lees-Edwards rule). */
int sc_rotate_by_gamma(sc_system* s, int k);
/* gradient_descent_minimize loop (src/cpp/relaxation_dynamics.cpp:33-85,
 * gradient_relax_step :11-31): at most nmax steps, stopping when E <= emin. */
int sc_gd_steps(sc_system* s, int nmax, double emin, int* steps_done);
/* Affine rescale of packing k (k = -1: all): x, y, L multiplied by scale[k]
 * (the quasi-static compression step of src/cuda/strain_dynamics.cu:1111-1128). */
int sc_affine_resize(sc_system* s, const double* scale);
/* The y-only form: y and Ly multiplied by scale[k], x and Lx kept
 * (resize_y_coords_2p and y_resize_step_2p, src/cuda/2point_dynamics.cu:1174-1295). */
int sc_affine_resize_y(sc_system* s, const double* scale);

/* FIRE minimisation of every packing to F_rms < ftol (no reference counterpart;
 * the reference's own minimisers are gradient descent, relaxation_dynamics.cpp,
 * and the GPU-side conjugate gradient of src/cuda/energy_minimize.cu:1196-1314,
 * whose F_rms convention :1277 is used here).  FIRE of Bitzek et al., PRL 97,
 * 170201 (2006), with the half-step retreat of FIRE 2.0 (Guenole et al., Comput.
 * Mater. Sci. 175, 109584 (2020)), in generalised coordinates (x, y,
 * psi = phi sqrt(I)) with generalised force g = (Fx, Fy, tau/sqrt(I)).  One
 * iteration per packing, all sums in FP64:
 *   1. forces at x; stop if sqrt(sum(Fx^2+Fy^2+tau^2)/(3N)) < ftol
 *   2. P = g.v, G = g.g, V = v.v
 *   3. P >= 0: v <- (1-alpha) v + alpha sqrt(V/G) g; npos += 1;
 *              if npos > n_min: dt <- min(f_inc dt, dt_max), alpha <- f_alpha alpha
 *      else:   x <- x - dt v / 2; v <- 0; dt <- f_dec dt; alpha <- alpha0; npos <- 0
 *   4. v <- v + dt g; displacement dt v with its (x, y) part shortened to
 *      max_move if longer; x <- x + displacement
 * params[10]: dt0 dt_max n_min f_inc f_dec alpha0 f_alpha max_move ftol max_iter.
 * In SC_PREC_MIXED mode the pair math runs in FP32 until F_rms < switch_frms,
 * then in FP64 to the end.  iters[npack] and frms[npack] may be NULL. */
int sc_fire_minimize(sc_system* s, const double* params, double switch_frms, int* iters,
                     double* frms);

/* Neighbour set of packing k under the reference's criterion (check_neighbors
 * src/cpp/spherocyl_box.cpp:303-318 over the own + 8 adjacent cells), as
 * half-list pairs (p < q) in caller particle indices.  *count gets the total;
 * only min(count, cap) pairs are written. */
int sc_get_neighbors(sc_system* s, int k, int* p, int* q, int cap, int* count);
/* Contacts (|D| < R_p + R_q) of packing k, p < q. */
int sc_get_contacts(sc_system* s, int k, int* p, int* q, int cap, int* count);

/* Per-kernel timing with HIP events on the system's stream. */
#define SC_KERNEL_PREDICT 0 /* Heun predictor force pass */
#define SC_KERNEL_CORRECT 1 /* Heun corrector force pass + cell binning */
#define SC_KERNEL_FORCE 2   /* plain force / stress pass */
#define SC_KERNEL_FIRE_FORCE 3
#define SC_KERNEL_FIRE_UPDATE 4
#define SC_KERNEL_SORT 5
#define SC_KERNEL_COUNT 6
int sc_timing_enable(sc_system* s, int enable);
int sc_timing_get(sc_system* s, int kernel, double* ms_total, long long* launches);
int sc_timing_reset(sc_system* s);
int sc_synchronize(sc_system* s);

/* Allocation-free calc_se.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The force / energy / stress pass
 * of sc_calc_se (list path, one lane per particle, 256-thread workgroups) on
 * one square packing whose arrays are already in device memory, in the caller's
 * particle order: x, y, phi, A, R (n each, device) at strain gamma (|gamma| <=
 * 1/2) in a box of side L with cell padding pad, contact model SC_CONTACT_*.
 * Writes fx, fy, tau (n each, device) and, when not NULL, se[10] (device):
 * E Pxx Pyy Pxy cPxx cPyy cPxy cPyx ncontacts F_rms as sc_get_box reports
 * them, and *status (device int): SC_OK, or SC_ERR_CELL_OVERFLOW /
 * SC_ERR_LIST_OVERFLOW / SC_ERR_ARG (non-finite coordinate) found on the
 * device.  All working memory is one static device array of
 * SC_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends): n above
 * SC_NOALLOC_MAX_PARTICLES is SC_ERR_ARG, and a cell grid above
 * SC_NOALLOC_MAX_CELLS cells is coarsened.  Launches on `stream` (a
 * hipStream_t; NULL: the null stream), returns without synchronising, and
 * calls no allocator.  Every call on one GPU shares that GPU's scratch array:
 * calls that overlap in time on one GPU race (serialise them on one stream). */
#define SC_NOALLOC_MAX_PARTICLES 32768
#define SC_NOALLOC_MAX_CELLS 32768
#define SC_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
int sc_calc_se_noalloc(int n, double L, double gamma, double pad, int contact_model,
                       const double* x, const double* y, const double* phi, const double* A,
                       const double* R, double* fx, double* fy, double* tau, double* se,
                       int* status, void* stream);

/* Seeded random placement (counterpart of SpherocylBox::random_configuration,
 * src/cpp/spherocyl_box.cpp:694-790, and its initialConfig modes,
 * src/cpp/spherocyl_box.h:8 plus the grid modes 4/5 of :718-737): centres
 * uniform in [0,L)^2 (sheared coordinates at strain gamma; grid modes: one per
 * site of a rows x cols grid in the particle aspect ratio) and angles uniform in
 * [0,pi) (ALIGNED modes: 0).  RANDOM modes redraw a particle whose segment comes
 * within sqrt(1e-5)(R_p+R_q) of an earlier one under the Lees-Edwards pair
 * separation; ZERO_E and GRID modes redraw any contact.  Host only; needs no
 * GPU.  SC_ERR_PLACEMENT after max_tries draws for one particle, or when the
 * grid sites are too small for the particles. */
#define SC_CONFIG_RANDOM 0
#define SC_CONFIG_RANDOM_ALIGNED 1
#define SC_CONFIG_ZERO_E 2
#define SC_CONFIG_ZERO_E_ALIGNED 3
#define SC_CONFIG_GRID 4
#define SC_CONFIG_GRID_ALIGNED 5
int sc_random_configuration(int n, double L, double gamma, const double* A, const double* R,
                            int config, unsigned long long seed, int max_tries, double* x,
                            double* y, double* phi);
/* Gamma-distributed aspect ratios: set_shapes with an aspect width > 0
 * (src/cuda/spherocyl_box.cpp:1266-1300).  A_p / R_p is drawn from a gamma
 * distribution of mean aspect_mean and standard deviation aspect_rel_sigma *
 * aspect_mean (Marsaglia-Tsang sampling of a seeded xoshiro256** stream, where
 * the reference uses an unseeded std::default_random_engine); R_p = 0.5
 * bidispersity for even p and 0.5 for odd p.  The reference then rescales the
 * box so that the drawn shapes fill the target packing fraction: size the box
 * from the drawn R and A (L = sqrt(sum R (pi R + 4 A) / phi)).  Host only. */
int sc_gamma_shapes(int n, double aspect_mean, double aspect_rel_sigma, double bidispersity,
                    unsigned long long seed, double* R, double* A);
/* sc_random_configuration in RANDOM mode. */
int sc_random_packing(int n, double L, double gamma, const double* A, const double* R,
                      unsigned long long seed, int max_tries, double* x, double* y, double* phi);

#ifdef __cplusplus
}
#endif
#endif /* SPHEROCYLS_H */
