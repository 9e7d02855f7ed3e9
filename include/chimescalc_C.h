/*
 * Per-cluster ChIMES C ABI (MI355X implementation).
 *
 * Same symbols and signatures as the reference's chimesFF/api/chimescalc_C.h:11-28,
 * with the same semantics (chimesFF/api/chimescalc_C.cpp:22-318): one static
 * evaluator set up by set_chimes / init_chimes / chimes_read_params; each
 * compute call adds its energy to *epot, its forces to force[][], and its
 * packed stress (xx xy xz yy yz zz) to stress[0..5] (the 4-body call copies
 * stress[0..8] back after the update, as the reference does).  Type names
 * are resolved per call; unknown names print an error and exit.  Each call
 * evaluates its cluster with the GPU kernels.
 */
#ifdef __cplusplus
extern "C" {
#endif

/* chimescalc_C.h:11-13 */
double get_chimes_max_2b_cutoff();
double get_chimes_max_3b_cutoff();
double get_chimes_max_4b_cutoff();
/* chimescalc_C.h:14-16 */
int get_chimes_2b_order();
int get_chimes_3b_order();
int get_chimes_4b_order();
/* chimescalc_C.h:17-21 */
void set_chimes();
void init_chimes(int rank);
void chimes_read_params(char *param_file);
void chimes_build_pair_int_trip_map();
void chimes_build_pair_int_quad_map();
/* chimescalc_C.h:22-24 */
void chimes_compute_2b_props(double rij, double dr[3], char *atype2b[2], double force[2][3], double stress[9],
                             double *epot);
void chimes_compute_3b_props(double dr_3b[3], double dist_3b[3][3], char *atype3b[3], double f3b[3][3],
                             double stress[9], double *epot);
void chimes_compute_4b_props(double dr_4b[6], double dist_4b[6][3], char *atype4b[4], double f4b[4][3],
                             double stress[9], double *epot);
/* chimescalc_C.h:26-28 */
void chimes_compute_2b_props_fromf90(double *rij, double dr[3], char *type1, char *type2, double force[2][3],
                                     double stress[9], double *epot);
void chimes_compute_3b_props_fromf90(double dr_3b[3], double dist_3b[3][3], char *type1, char *type2, char *type3,
                                     double f3b[3][3], double stress[9], double *epot);
void chimes_compute_4b_props_fromf90(double dr_4b[6], double dist_4b[6][3], char *type1, char *type2, char *type3,
                                     char *type4, double f4b[4][3], double stress[9], double *epot);

#ifdef __cplusplus
}
#endif
