/* wipems_nf2ff.h -- near-field to far-field transform on the GPU.
 *
 * Replaces nf2ff_calc (nf2ff/nf2ff_calc.cpp) for Cartesian meshes:
 *
 *   nf2ff_create     nf2ff_calc::nf2ff_calc (:164-214): frequency, theta and
 *                    phi lists (rad), centre (m), far-field radius (m, the
 *                    reference's default is 1), relative eps/mu of the medium
 *   nf2ff_add_plane  nf2ff_calc::AddPlane / AddSinglePlane (:330-528): one
 *                    face of frequency-domain E and H at its mesh nodes; one
 *                    axis has a single line (the normal), the other two at
 *                    least three
 *   nf2ff_get        GetETheta/GetEPhi/GetHTheta/GetHPhi/GetRadPower/
 *                    GetTotalRadPower/GetMaxDirectivity (nf2ff_calc.h:60-72)
 *
 * Per face the device evaluates, for every (theta, phi), the area-weighted
 * sums of n x H and -n x E times exp(jk r.rhat) over the face nodes (one
 * workgroup per direction, FP64 phases and accumulators); the host combines
 * them into N_theta, N_phi, L_theta, L_phi and the far fields in the
 * reference's formulas.  The reference evaluates the per-node terms in float;
 * this computes them in double.
 *
 * Errors: negative FDTD_ERR_* codes, message via fdtd_last_error().
 */
#ifndef WIPEMS_NF2FF_H
#define WIPEMS_NF2FF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nf2ff_calc nf2ff_calc;

int nf2ff_create(double freq, const double* theta, uint32_t n_theta, const double* phi, uint32_t n_phi,
                 const double center[3], double radius, double eps_r, double mue_r, nf2ff_calc** out);

/* nf2ff_calc::SetMirror (:231-245): a symmetry plane at `pos` (m) normal to
 * `axis`, type 0 off, 1 PEC, 2 PMC.  Later nf2ff_add_plane calls add each
 * face's images as AddPlane does (:281-328): one plane, two planes (three
 * images) or all three (seven). */
int nf2ff_set_mirror(nf2ff_calc* c, int type, int axis, double pos);

/* lines_*: node coordinates in metres; E, H: [3][nx][ny][nz] complex values as
 * (re, im) double pairs (the field-box accumulators, fdtd_read_dft). */
int nf2ff_add_plane(nf2ff_calc* c, const double* lines_x, uint32_t nx, const double* lines_y, uint32_t ny,
                    const double* lines_z, uint32_t nz, const double* E, const double* H);

/* Far fields [n_theta][n_phi] as (re, im) pairs, radiated power density
 * [n_theta][n_phi], total radiated power through the faces, maximum
 * directivity.  Any output pointer may be NULL. */
int nf2ff_get(nf2ff_calc* c, double* E_theta, double* E_phi, double* H_theta, double* H_phi, double* P_rad,
              double* rad_power, double* max_dir);

int nf2ff_destroy(nf2ff_calc* c);

#ifdef __cplusplus
}
#endif

#endif
