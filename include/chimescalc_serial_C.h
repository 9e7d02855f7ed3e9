/*
 * System-level ChIMES C ABI (MI355X implementation).
 *
 * Same symbols, signatures and semantics as the reference's
 * serial_interface/api/chimescalc_serial_C.h:11-18, so C, Fortran (bind(C))
 * and Python ctypes callers link against this library unchanged:
 *   - a static default instance (set_chimes_serial / init_chimes_serial /
 *     calculate_chimes) and handle-based instances (chimescalc_serial_C.cpp:45-124);
 *   - energy and forces are accumulated onto the caller's values, the energy
 *     (seed included) is divided by (R+1)^3 when replicating, stress[9] is
 *     overwritten with virial / volume in kcal/mol/A^3;
 *   - init_chimes_serial_instance ignores rank (chimescalc_serial_C.cpp:85-90);
 *   - errors print a message and exit, as in the reference.
 * The evaluation runs on the GPU.
 */
#ifdef __cplusplus
extern "C" {
#endif

/* chimescalc_serial_C.h:11 */
void *chimes_open_instance();
/* chimescalc_serial_C.h:12 */
void chimes_close_instance(void *handle);
/* chimescalc_serial_C.h:13 -- small = 1 replicates cells shorter than the 2B cutoff */
void set_chimes_serial(int small);
/* chimescalc_serial_C.h:14 */
void set_chimes_serial_instance(void *handle, int small);
/* chimescalc_serial_C.h:15 */
void init_chimes_serial(char *param_file, int *rank);
/* chimescalc_serial_C.h:16 */
void init_chimes_serial_instance(void *handle, char *param_file, int rank);
/* chimescalc_serial_C.h:17 */
void calculate_chimes(int natom, double *xc, double *yc, double *zc, char *atom_types[], double ca[3],
                      double cb[3], double cc[3], double *energy, double fx[], double fy[], double fz[],
                      double stress[9]);
/* chimescalc_serial_C.h:18 */
void calculate_chimes_instance(void *handle, int natom, double *xc, double *yc, double *zc, char *atom_types[],
                               double ca[3], double cb[3], double cc[3], double *energy, double fx[], double fy[],
                               double fz[], double stress[9]);

#ifdef __cplusplus
}
#endif
