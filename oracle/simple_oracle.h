/*
 * CPU oracle for the Ixchel2D SIMPLE iteration (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * upstream algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product (libixchel2d.so) never links
 * or calls it.
 *
 * Arrays use the upstream Fortran layout: element (i,j), 1-based, of an array
 * with first extent m sits at (i-1) + (j-1)*m.  Extents: u(mi,nj+1),
 * v(mi+1,nj), T/p/p'/b_o(mi+1,nj+1), au(mi,nj+1), av(mi+1,nj).
 */
#ifndef IXCHEL2D_SIMPLE_ORACLE_H
#define IXCHEL2D_SIMPLE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ox_state ox_state;

/* per-iteration record: v error, p' error, maxbo, T error, u residual norm,
 * the pass counts of the momentum, p' and energy loops, the SIMPLE counter
 * after the exit test and whether upstream writes the line to fort.102 */
typedef struct {
  double err_v, err_p, maxbo, err_t, residuo;
  int32_t passes_mom, passes_p, passes_t, simple_iter, logged, pad_;
} ox_hist;

ox_state *ox_create(int32_t mi, int32_t nj, const double *xu, const double *yv,
                    const double *xp, const double *yp);
void ox_destroy(ox_state *s);

/* scalars: dt, rel_pres, rel_vel, rel_ener, gamma_momen, gamma_energ, Ri, Riy,
 * conv_u, conv_t, conv_p, conv_resi, conv_paso */
void ox_set_params(ox_state *s, const double *scalars13, int32_t iter_ecuaci_max,
                   int32_t iter_simple_max);

/* var: 0=u 1=v 2=T; side: 0..3 = a,b,c,d.  type[k]: 0 none, 1 Dirichlet,
 * 2 Neumann, for node k+1 along the side (length of the side's extent). */
void ox_set_bc(ox_state *s, int32_t var, int32_t side, const int32_t *type, const double *value);
/* var: 0=u 1=v 2=T; Sc, Sp in the variable's own extent (NULL = zero) */
void ox_set_sources(ox_state *s, int32_t var, const double *sc, const double *sp);

/* field ids */
enum { OX_U = 0, OX_V, OX_T, OX_P, OX_U_ANT, OX_V_ANT, OX_T_ANT, OX_PCORR, OX_BO,
       OX_AU, OX_AV, OX_RESU, OX_NFIELDS };
void ox_set_field(ox_state *s, int32_t field, const double *src);
void ox_get_field(const ox_state *s, int32_t field, double *dst);
int64_t ox_field_size(const ox_state *s, int32_t field);

/* phases of one SIMPLE iteration (source/main.f90:404-1270) */
void ox_momentum_u_y(ox_state *s);
void ox_momentum_u_x(ox_state *s);
void ox_momentum_v_y(ox_state *s);
void ox_momentum_v_x(ox_state *s);
void ox_pcorr_zero(ox_state *s);
void ox_pcorr_y(ox_state *s);
void ox_pcorr_x(ox_state *s);
void ox_correct(ox_state *s);
void ox_energy_y(ox_state *s);
void ox_energy_x(ox_state *s);
double ox_residual_u(ox_state *s);
void ox_lag(ox_state *s);

/* whole time steps with upstream's convergence control; writes one record
 * per SIMPLE iteration into hist (capacity hist_cap), returns the count */
int32_t ox_run(ox_state *s, int32_t nsteps, ox_hist *hist, int32_t hist_cap);

/* one Thomas solve, in place (source/mod_solucionador.f90:25-47) */
void ox_tridiagonal(const double *a, double *b, const double *c, double *r, int32_t n);

/* Nusselt integrals on the horizontal walls (source/mod_postproceso.f90:100-215) */
void ox_nusselt_y(const ox_state *s, int32_t i0, int32_t i1, double *nu0, double *nu1);
/* mean Nusselt on the vertical walls x=0 (hot) and x=W (cold) */
void ox_nusselt_x(const ox_state *s, double *nu_west, double *nu_east);

#ifdef __cplusplus
}
#endif
#endif
