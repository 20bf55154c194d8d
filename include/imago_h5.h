/* imago_h5.h - the reference's SCF integral file (HDF5), for checkpoint and
 * resume of the integral stage and for exchanging integrals with the
 * reference's Fortran host.  libimago_h5.so (host code over the HDF5 C
 * library; the GPU library does not depend on it).
 *
 * The layout is initHDF5_SCF / initSCFIntegralHDF5's
 * (src/imago/hdf5SCF.F90:60-230, src/imago/hdf5SCFIntg.F90:105-525): a
 * group per k-point set, under it atomIntgGroup with the groups
 * atomOverlap, atomKEOverlap, atomMVOverlap, atomNPOverlap,
 * atomDMOverlap, atomMMOverlap, atomKOverlap, atomKOverlapPlusG and
 * atomPotOverlap (one subgroup per potential term), a packed dataset per
 * k-point ("0000001" ...), and a "status" attribute on every group that
 * the reference's integral routines read to skip finished work
 * (src/imago/integrals.F90:265-294).  All functions return IMAGO_H5_OK or
 * IMAGO_H5_E (message from imago_h5_last_error). */
#ifndef IMAGO_H5_H
#define IMAGO_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define IMAGO_H5_OK 0
#define IMAGO_H5_E (-1)

/* the matrices of the file; OL, KE, NP and EP take the values of
 * IMAGO_OL .. IMAGO_EP of imago_hip.h */
#define IMAGO_H5_OL 1
#define IMAGO_H5_KE 2
#define IMAGO_H5_NP 3
#define IMAGO_H5_EP 4
#define IMAGO_H5_MV 5

typedef struct imago_h5 imago_h5;

/* Open the file at path (created when absent) and its k-point-set group
 * kpoint_group (created with every group, dataset and status attribute, all
 * statuses 0, when absent: the reference names it after the k-point mesh,
 * e.g. "00004_00004_00004").  *created = 1 when the group was made. */
int imago_h5_open(const char *path, const char *kpoint_group, int num_kpoints, int vale_dim,
                  int pot_dim, int gamma, imago_h5 **out, int *created);
int imago_h5_close(imago_h5 *f);

/* One packed matrix (imago_get_packed's layout: 2 n(n+1)/2 doubles, real
 * and imaginary interleaved, or n(n+1)/2 for gamma) of kind, potential
 * term (EP only, else 0), k-point (0-based). */
int imago_h5_write_packed(imago_h5 *f, int kind, int term, int kpoint, const double *packed);
int imago_h5_read_packed(imago_h5 *f, int kind, int term, int kpoint, double *packed);

/* The group's status attribute: 1 once all of its k-points are written. */
int imago_h5_status(imago_h5 *f, int kind, int term, int *status);
int imago_h5_set_status(imago_h5 *f, int kind, int term, int status);

const char *imago_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
