/*
 * petar_host.h -- C ABI of libpetar_host.so: the host side that feeds the
 * soft-force path.
 *
 * PeTar gets its interaction lists from FDPS (TreeForForceLong<...>::
 * QuadrupoleWithSymmetrySearch, src/petar.hpp:619-623, driven at
 * src/petar.hpp:894-899), which is not part of the reference repository.
 * This library is the stand-alone replacement used to drive the path:
 *   - initial conditions (the Plummer recipe of
 *     src/particle_distribution_generator.hpp:173-250 with its MT19937
 *     stream, plus Kroupa-IMF masses, King W0 models and primordial
 *     binaries);
 *   - an octree with FDPS's parameters (theta, n_leaf_limit,
 *     n_group_limit), quadrupole moments about the centre of mass, i-groups
 *     and the per-group EPJ/SPJ interaction lists with symmetric r_search
 *     inclusion.
 * All functions return 0 on success and a negative code on error; the text
 * of the last error of the calling thread is in pth_host_last_error().
 */
#ifndef PETAR_HOST_H
#define PETAR_HOST_H
#ifdef __cplusplus
extern "C" {
#endif

const char *pth_host_last_error(void);

/* ---- MT19937 (the generator FDPS's MTTS wraps) ---------------------------- */
typedef struct pth_mt pth_mt;
pth_mt *pth_mt_new(unsigned int seed);          /* init_genrand(seed) */
void pth_mt_free(pth_mt *mt);
unsigned int pth_mt_int32(pth_mt *mt);          /* genrand_int32 */
double pth_mt_res53(pth_mt *mt);                /* genrand_res53 */
double pth_mt_real2(pth_mt *mt);                /* genrand_real2, [0,1) */

/* ---- initial conditions ---------------------------------------------------- */
/* Equal-mass Plummer exactly as makePlummerModel(mass_glb, n, n, ..., eng,
 * seed) (src/particle_distribution_generator.hpp:173-250); vel may be NULL. */
int pth_ic_plummer_equal(long n, double mass_glb, double eng, unsigned int seed,
                         double *mass, double *pos, double *vel);

/* Kroupa (2001) IMF masses in [m_lo, m_hi] (solar units, slopes 1.3 below
 * 0.5 Msun and 2.3 above), scaled so that they sum to m_tot (m_tot <= 0:
 * left in solar units). */
int pth_ic_kroupa_masses(long n, double m_lo, double m_hi, double m_tot, unsigned int seed,
                         double *mass);

/* Cluster of n particles: Plummer (model 0) or King with W0 (model 1),
 * Kroupa masses (kroupa != 0) or equal masses, a fraction f_bin of the
 * particles paired into binaries (log-uniform semi-major axis in
 * [a_lo, a_hi], thermal eccentricity, isotropic orientation, uniform mean
 * anomaly), N-body units (G = 1, M = 1, E = -1/4 for the centre-of-mass
 * system).  segregate != 0 assigns masses to radii in rank order (heaviest
 * innermost).  vel may be NULL. */
int pth_ic_cluster(long n, int model, double w0, int kroupa, double f_bin, double a_lo,
                   double a_hi, int segregate, unsigned int seed,
                   double *mass, double *pos, double *vel);

/* ---- octree + group walk --------------------------------------------------- */
typedef struct pth_tree pth_tree;

/* Builds the tree over n particles (positions FP64 [n][3], masses, r_search).
 * nthreads <= 0 uses the OpenMP default. */
int pth_tree_build(pth_tree **tree, long n, const double *pos, const double *mass,
                   const double *r_search, int n_leaf_limit, int n_group_limit, int nthreads);

/* Builds the i-groups and their interaction lists with opening angle theta
 * (theta <= 0 opens every cell: every j is an EPJ). */
int pth_tree_walk(pth_tree *tree, double theta);

/* Sizes: tree nodes (= SPJ records), groups, total EPJ and SPJ list entries. */
int pth_tree_sizes(const pth_tree *tree, long *n_node, long *n_group, long *n_epj_list,
                   long *n_spj_list);

/* order[k] = input index of the k-th particle in tree (Morton) order. */
int pth_tree_order(const pth_tree *tree, long *order);

/* SPJ records = all tree nodes: mass, centre of mass, raw second moment
 * sum m (x-c)(x-c)^T as xx, yy, zz, xy, xz, yz; plus per node the particle
 * range [begin, end) in tree order and the level. Any pointer may be NULL. */
int pth_tree_nodes(const pth_tree *tree, double *mass, double *com, double *quad,
                   long *begin, long *end, int *level);

/* Groups: particle range [begin, begin+n_epi) in tree order and list sizes. */
int pth_tree_groups(const pth_tree *tree, long *epi_begin, int *n_epi, int *n_epj, int *n_spj);

/* Concatenated lists in group order: EPJ entries are tree-order particle
 * indices, SPJ entries are node indices. */
int pth_tree_lists(const pth_tree *tree, int *id_epj, int *id_spj);

void pth_tree_free(pth_tree *tree);

/* ---- PeTar snapshot files ------------------------------------------------ */
/* The files PeTar writes and reads with writeParticleAscii / Binary
 * (src/petar.hpp:1905-1908, 2757-2763): a FileHeader (src/io.hpp:210-282:
 * file id, N, time [, pos_offset[3], vel_offset[3]]) and one FPSoft record
 * per particle (ParticleBase, Ptcl, FPSoft fields, src/particle_base.hpp,
 * src/ptcl.hpp, src/soft_ptcl.hpp).  Columns, by the names of PeTar's own
 * reader (tools/analysis/data.py:390-399): mass, pos[3], vel[3],
 * binary_state, [radius, dm, time_record, time_interrupt], r_search, id,
 * mass_bk, status, r_in, r_out, acc_soft[3], pot, pot_soft, [pot_ext], n_nb;
 * binary_state, id, mass_bk, status and n_nb are int64, the others FP64. */
#define PTH_SNAP_BINARY 0   /* data_format r/w BINARY (petar -i 0) */
#define PTH_SNAP_ASCII 1    /* data_format r/w ASCII (petar -i 1, the default) */
#define PTH_SNAP_EXTERNAL 1 /* layout flag: external-potential build (header offsets, pot_ext) */
#define PTH_SNAP_SE 2       /* layout flag: STELLAR_EVOLUTION base fields (no BSE star data) */
typedef struct pth_snapshot pth_snapshot;
/* Reads a snapshot of the given format and layout flags; a file whose size
 * or records do not match the layout is an error naming the mismatch. */
int pth_snapshot_read(const char *path, int format, int flags, pth_snapshot **snap);
/* An empty (zeroed) snapshot of n particles, for writing. */
int pth_snapshot_new(long long n, int flags, pth_snapshot **snap);
int pth_snapshot_header(const pth_snapshot *snap, long long *file_id, long long *n, double *time,
                        double *pos_offset, double *vel_offset);
int pth_snapshot_set_header(pth_snapshot *snap, long long file_id, double time, const double *pos_offset,
                            const double *vel_offset);
/* Width (1 or 3) of a column, 0 when the layout has none of that name. */
int pth_snapshot_column_width(const pth_snapshot *snap, const char *name);
/* Copies a column out ([n] or [n][3], int64 or FP64 as listed above) / in. */
int pth_snapshot_column(const pth_snapshot *snap, const char *name, void *out);
int pth_snapshot_set_column(pth_snapshot *snap, const char *name, const void *in);
int pth_snapshot_write(const pth_snapshot *snap, const char *path, int format);
void pth_snapshot_free(pth_snapshot *snap);

#ifdef __cplusplus
}
#endif
#endif
