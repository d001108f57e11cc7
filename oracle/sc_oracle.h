// Test infrastructure only: the CPU oracle restatement.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// to test results against or as the timed CPU baseline -- never on the product
// path.
//
// C ABI of liboracle.so.  Every function follows the reference CPU code
// (tmarschall/spherocyls src/cpp) cited beside it; FIRE has no reference
// counterpart and is the project's own specification (see sc_oracle.cpp).
#ifndef SC_ORACLE_H
#define SC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_box orc_box;

// SpherocylBox(dL, N, x, y, phi, A*, R*, pad)  src/cpp/spherocyl_box.cpp:152-173
orc_box* orc_new(double L, int n, const double* x, const double* y, const double* phi,
                 const double* A, const double* R, double pad);
void orc_free(orc_box* b);
void orc_set_params(orc_box* b, double gamma, double gamma_total, double strain_rate,
                    double resize_rate, double step);
// Overwrite positions without wrapping (tests that probe images / invariances).
void orc_set_positions(orc_box* b, const double* x, const double* y, const double* phi);
// Overwrite the box lengths (states read from non-square files); rebuilds cells.
void orc_set_box(orc_box* b, double Lx, double Ly);

void orc_find_neighbors(orc_box* b);   // spherocyl_box.cpp:320-355
void orc_calc_forces(orc_box* b);      // strain_dynamics.cpp:137-153
void orc_calc_se(orc_box* b);          // spherocyl_box.cpp:502-522
void orc_strain_step(orc_box* b);      // strain_dynamics.cpp:179-235
int orc_run_steps(orc_box* b, int nsteps, int se_every, double* se_out);  // :285-296
void orc_resize_step(orc_box* b, int shrink, int corrected_image);  // resize_dynamics.cpp:164-243
void orc_gd_step(orc_box* b);          // relaxation_dynamics.cpp:11-31
void orc_pure_step(orc_box* b);        // src/cuda/2point_dynamics.cu:1663-2285, CPU conventions
void orc_flip_shear(orc_box* b);       // src/cuda/find_neighbors.cu:552-579
void orc_rotate_by_gamma(orc_box* b);  // src/cuda/find_neighbors.cu:580-645
void orc_set_back_gamma(orc_box* b);   // spherocyl_box.cpp:524-545
// 0: closest-point contact (src/cpp); 1: two-point contact model
// (src/cuda/2point_dynamics.cu:391-546), each pair evaluated once, lower id first.
void orc_set_contact_model(orc_box* b, int model);

// Per-particle state, reference particle order.
void orc_get_particles(const orc_box* b, double* x, double* y, double* phi, double* R, double* A,
                       double* fx, double* fy, double* tau, double* I, double* RC);
// Sum over contacts of |f_ij| components per particle (scale for tolerances).
void orc_force_scale(orc_box* b, double* sx, double* sy, double* st);
// out[0..19]: Lx Ly gamma gamma_total strain_rate resize_rate step packing pad
//   E Pxx Pyy Pxy (reference formulas) cPxx cPyy cPxy cPyx (corrected virial)
//   rows cols ncontacts
void orc_get_scalars(const orc_box* b, double* out);
int orc_get_neighbors(const orc_box* b, int* p, int* q, int cap);
// Contacts found through the neighbour list / by brute force over all pairs.
int orc_get_contacts(orc_box* b, int* p, int* q, int cap);
int orc_brute_contacts(orc_box* b, int* p, int* q, int cap);

// FIRE minimisation (project specification, FP64).  params[0..9]:
//   dt0 dt_max n_min f_inc f_dec alpha0 f_alpha max_move ftol max_iter
// Returns iterations used; *frms_out gets the final residual.
int orc_fire(orc_box* b, const double* params, double* frms_out);
// One FIRE iteration from an explicit (v, dt, alpha, npos) state.
// fire_state[0..3] = dt alpha npos converged (updated in place).
void orc_fire_iter(orc_box* b, const double* params, double* vx, double* vy, double* vpsi,
                   double* fire_state);
// One pair's contacts (model 0: closest point, 1: two-point) in an open box of
// side L: out[9 k ..] = s t Dx Dy dist fx fy tau_p tau_q; returns the count.
int orc_pair_contacts(int model, double xp, double yp, double php, double Rp, double Ap,
                      double xq, double yq, double phq, double Rq, double Aq, double L,
                      double gamma, double* out);
double orc_frms(orc_box* b);  // sqrt(sum(Fx^2+Fy^2+tau^2)/(3N)) at the current state

#ifdef __cplusplus
}
#endif
#endif
