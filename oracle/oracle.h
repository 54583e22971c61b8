/*
 * oracle.h -- FP64 CPU restatement of PeTar's soft-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (petar_amd/, libpetar_hip.so) never links or calls it.
 *
 * Conventions shared by every function:
 *   - positions are [n][3] row-major FP64, quadrupoles are [n][6] in the
 *     order xx, yy, zz, xy, xz, yz (FDPS F64mat member names);
 *   - a "walk" is one FDPS i-group with its interaction lists: walk w owns
 *     n_epi[w] consecutive entries of epi_pos/epi_rs and n_epj[w] / n_spj[w]
 *     consecutive entries of id_epj / id_spj (indices into epj_* / spj_*);
 *   - acc/pot/nnb are ASSIGNED (not accumulated), as RetrieveForceCUDA does
 *     (src/force_gpu_cuda.cu:866-875).
 *
 * Neighbour-count conventions (nnb_mode), see src/force_gpu_cuda.cu:62-64,
 * src/phantomquad_for_p3t_x86.hpp:516-525 and src/soft_force.hpp:57-64:
 *   0  GPU reference:  r^2 + eps^2 <  max(rs_i, rs_j)^2
 *   1  x86 SIMD:       r^2 + eps^2 <= max(rs_i, rs_j)^2
 *   2  scalar FP64:    r^2         <  max(rs_i, rs_j)^2
 *
 * Parity status of each function is recorded in the header comment of its
 * definition in oracle.c.
 */
#ifndef PETAR_ORACLE_H
#define PETAR_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* EPJ (linear cutoff) + SPJ force over a set of walks: quadrupole SPJ
 * (src/force_gpu_cuda.cu:276-323, src/soft_force.hpp:160-200), or with
 * spj_quad == NULL the monopole SPJ (src/soft_force.hpp:125-158). */
int orc_soft_force_walks(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                         const double *epi_pos, const double *epi_rs,
                         const int *id_epj, const int *id_spj,
                         const double *epj_pos, const double *epj_mass, const double *epj_rs,
                         const double *spj_pos, const double *spj_mass, const double *spj_quad,
                         double eps2, double rout2, double G, int nnb_mode,
                         double *acc, double *pot, long long *nnb);

/* EPJ part only (acc/pot/nnb of the particle-particle list). */
int orc_epj_walks(int n_walk, const int *n_epi, const int *n_epj,
                  const double *epi_pos, const double *epi_rs, const int *id_epj,
                  const double *epj_pos, const double *epj_mass, const double *epj_rs,
                  double eps2, double rout2, double G, int nnb_mode,
                  double *acc, double *pot, long long *nnb);

/* SPJ part only (acc/pot of the superparticle list; spj_quad == NULL:
 * monopole). */
int orc_spj_walks(int n_walk, const int *n_epi, const int *n_spj,
                  const double *epi_pos, const int *id_spj,
                  const double *spj_pos, const double *spj_mass, const double *spj_quad,
                  double eps2, double G, double *acc, double *pot);

/* Number of EPJ pairs per i whose neighbour predicate lies inside the FP32
 * rounding band |r2+eps2 - rs2| <= rel * (rs2 + r2 + 2 r (|xi-c| + |xj-c|)),
 * c = centres[w] (pass NULL for absolute coordinates, c = 0). */
int orc_nnb_ambiguous_walks(int n_walk, const int *n_epi, const int *n_epj,
                            const double *epi_pos, const double *epi_rs, const int *id_epj,
                            const double *epj_pos, const double *epj_rs,
                            const double *centres, double eps2, double rel,
                            long long *n_amb);

/* All-pairs linear-cutoff force on sampled i (self pair included, as in a
 * tree list that contains the particle itself). */
int orc_direct_linear_cutoff(int n, const double *pos, const double *mass, const double *rs,
                             int n_sample, const int *idx_sample,
                             double eps2, double rout2, double G, int nnb_mode,
                             double *acc, double *pot, long long *nnb);

/* All-pairs Newtonian (no cutoff) force on sampled i, self pair skipped. */
int orc_direct_newton(int n, const double *pos, const double *mass,
                      int n_sample, const int *idx_sample, double eps2, double G,
                      double *acc, double *pot);

/* Mass, centre of mass and raw second moment sum m (x-c)(x-c)^T. */
void orc_moments(int n, const double *pos, const double *mass,
                 double *m_out, double *com_out, double *quad_out);

/* Changeover function (src/changeover.hpp). */
void orc_changeover_setr(double m_fac, double r_in, double r_out,
                         double *r_in_out, double *r_out_out);
double orc_changeover_wpot(double r, double r_in, double r_out);
double orc_changeover_wacc(double r, double r_in, double r_out);
double orc_changeover_wacc1(double r, double rdot, double r_in, double r_out);
double orc_changeover_wpot_two(double r, double rin1, double rout1, double rin2, double rout2);
double orc_changeover_wacc_two(double r, double rin1, double rout1, double rin2, double rout2);

/* Changeover correction of one i against its neighbour set (self excluded
 * by id), FP64 branch of src/hard.hpp:1404-1476 plus the self-potential term
 * of src/hard.hpp:1669-1673.  grp_i [2] and grp_j [n_ngb][2] are the
 * particles' group data (mass_backup, status; NULL: singles), which select
 * the single / member / artificial rules of src/hard.hpp:1453-1472 and the
 * self-term condition.  Adds to acc[3] and to pot_soft / pot_tot. */
void orc_changeover_correct_one(const double *pos_i, double mass_i, double rin_i, double rout_i,
                                long long id_i, int n_ngb, const double *pos_j,
                                const double *mass_j, const double *rin_j, const double *rout_j,
                                const long long *id_j, double eps2, double r_out_glb, double G,
                                double *acc, double *pot_soft, double *pot_tot, const double *grp_i,
                                const double *grp_j);

/* Changeover mode (the composition src/hard.hpp:1404-1476,1655-1691 applies
 * on top of the linear-cutoff force) over walks; see oracle.c. */
int orc_soft_force_walks_co(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                            const double *epi_pos, const double *epi_rs, const long long *epi_id,
                            const int *id_epj, const int *id_spj,
                            const double *epj_pos, const double *epj_mass, const double *epj_rs,
                            const long long *epj_id, const double *epj_rin, const double *epj_rout,
                            const double *spj_pos, const double *spj_mass, const double *spj_quad,
                            double eps2, double rout2, double G, double ngb_factor, double amb_rel,
                            double *acc, double *pot_soft, double *pot_tot, long long *nnb,
                            long long *n_amb, const double *epj_group);

/* Neighbour lists (FP64, r^2 < max(rs_i, rs_j)^2, self included) over each
 * walk's EPJ list: counts, and with ngb != NULL the entries at offset[i]. */
long long orc_ngb_walks(int n_walk, const int *n_epi, const int *n_epj, const double *epi_pos,
                        const double *epi_rs, const int *id_epj, const double *epj_pos,
                        const double *epj_rs, long long *count, const long long *offset, int *ngb);

#ifdef __cplusplus
}
#endif
#endif
