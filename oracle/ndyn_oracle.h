/*
 * ndyn_oracle.h -- CPU FP64 restatement of Ndyn's MD step (TEST INFRASTRUCTURE).
 *
 * This library is the parity checker for the HIP engine in ndyn_omp_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to check or time against it; the product path never
 * calls it.
 *
 * Conventions follow the reference (Ndyn, src/ of stewmasterj/Ndyn_OMP):
 *   - positions are scaled coordinates, AoS pos[3*i+a], nominally in [-0.5,0.5)
 *     and never wrapped (src/nd.f90:1029-1046);
 *   - types are 1-based, pairmat[(ti-1) + (tj-1)*ntypes] holds 1-based pair ids
 *     with 0 meaning "no interaction" (src/mods.f90:54, src/read_input.f90:127-133);
 *   - tables are (ptsize+1) x npairs column-major (src/potentials.f90:17);
 *   - atom indices in the returned half list are 0-based.
 * Units: eV, Angstrom, amu; time unit 10.1805057 fs.
 */
#ifndef NDYN_ORACLE_H
#define NDYN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* pair-potential kinds accepted by ndo_pair_table */
enum {
  NDO_MORSE = 1,      /* reference Morse: table force = -V'(r)   (src/potentials.f90:184-204) */
  NDO_BUCK = 2,       /* Buckingham A exp(-r/rho) - C/r^6        (src/potentials.f90:112-122,206-216) */
  NDO_LJ = 3,         /* Lennard-Jones 4 eps((s/r)^12-(s/r)^6)   (src/potentials.f90:124-133,218-227) */
  NDO_MORSE_PHYS = 4, /* energy-consistent Morse: table force = -r V'(r) (not in the reference) */
  NDO_SPRING = 5,     /* 'spring' (src/potentials.f90:88-92,178-183): a table for output only;
                         compute2/3 take the force from the strain against ipos */
  NDO_SINE = 6, NDO_COSINE = 7, NDO_SOFT = 8,  /* src/potentials.f90:82-87, 172-177 */
  NDO_JOHNSON = 9,                              /* 'John' :136-147, :230-241 */
  NDO_MORSE_C = 10, NDO_BUCK_C = 11, NDO_LJ_C = 12  /* + damped shifted Coulomb term */
};
/* added to a kind (not in the reference): energy table V(r) - V(rc), force table unchanged */
enum { NDO_ESHIFT = 0x100 };

/* list-rebuild trigger */
enum {
  NDO_TRIGGER_COMPAT = 0, /* reference: top-2 squared displacement in scaled units vs skin^2 in A^2 */
  NDO_TRIGGER_SKIN = 1    /* displacement in A: d1 + d2 + box-strain term > skin, re-grid on rebuild */
};

/* error codes (map the reference's STOP messages) */
enum {
  NDO_OK = 0,
  NDO_ERR_NEIMAX = 1,   /* 'too many neighbours in list'   src/nd.f90:1378,1421 */
  NDO_ERR_OUTSIDE = 2,  /* 'Particle outside of box'       src/nd.f90:1494 */
  NDO_ERR_ARG = 4
};

/* Build one pair table: P[i] = V(r_i), F[i] = r_i * dpot(r_i), r_i = sqrt(ptrmin + i*dr2),
 * i = 0..ptsize (src/potentials.f90:19-34). params = {rc, p2, p3, p4} in deck order. */
int ndo_pair_table(int kind, const double params[4], double ptrmin, int ptsize,
                   double *P, double *F, double *rc2, double *dr2);
/* the same with the pair's charge product qprd (Coulomb kinds) */
int ndo_pair_table_q(int kind, const double params[4], double qprd, double ptrmin, int ptsize,
                     double *P, double *F, double *rc2, double *dr2);

/* Reference NeiMax default (src/nd.f90:84-88). */
int ndo_neimax_estimate(double rcut, double skin, int np, const double box[3]);

/* Cell grid MCELL = int(box/(rcut+skin)) (src/nd.f90:80-83). */
void ndo_cell_grid(const double box[3], double rcut, double skin, int mcell[3]);

/* Half Verlet list (src/nd.f90:1214-1446).  When every mcell >= 3 this follows the
 * reference's linked-cell walk and entry order exactly; otherwise (brute != 0 or any
 * mcell < 3) it uses an O(N^2) minimum-image walk with i<j, the pair set the cell walk
 * would give on a large enough grid.  Output: vnl[(neimax+1)*i + 0] = count,
 * vnl[(neimax+1)*i + 1 + k] = 0-based j.  Also writes the cell of each atom when
 * cell_of != NULL.  Returns NDO_ERR_NEIMAX on overflow. */
int ndo_build_half_list(int np, const double *pos, const int *typ, const double box[3],
                        const int mcell[3], int ntypes, const int *pairmat,
                        const double *rc_pair, double skin, double ptrmin, int neimax,
                        int brute, int *vnl, int *cell_of);

/* compute0 (src/nd.f90:200-363) on a half list.  force[3*np] is overwritten;
 * pe / virial[6] are written only when cu / cv.  pene (optional) receives the
 * per-atom energy e_i = 1/2 sum_j V_ij when cu (the reference leaves it zero). */
void ndo_compute0(int np, const double *pos, const int *typ, const double box[3],
                  const int *vnl, int neimax, int ntypes, const int *pairmat,
                  const double *P, const double *F, const double *rc2, const double *dr2,
                  double ptrmin, int ptsize, int cv, int cu,
                  double *force, double *pe, double virial[6], double *pene);
/* compute1 (asymmetric pair matrices)  src/nd.f90:364-530 */
void ndo_compute1(int np, const double *pos, const int *typ, const double box[3],
                  const int *vnl, int neimax, int ntypes, const int *pairmat,
                  const double *P, const double *F, const double *rc2, const double *dr2,
                  double ptrmin, int ptsize, int cv, int cu,
                  double *force, double *pe, double virial[6], double *pene);
int ndo_asymmetric(int ntypes, const int *pairmat);
/* integrate1 with held atoms (src/nd.f90:1051-1075): an atom of a hold group with a
 * displacement condition moves by bc_parm along the constrained axes. */
void ndo_integrate1_hold(int np, double *pos, const double *vel, const double *acc,
                         double *disp, const double box[3], double dt, const int *hold,
                         const int *bc_disp, const int *bc_dir, const double *bc_parm);

/* compute2 (src/nd.f90:531-727, asym = 0) and compute3 (src/nd.f90:728-933, asym = 1):
 * compute0 / compute1 where the pairs flagged in spring[] (npairs entries) take their force
 * from the strain s = r/r0 - 1 against the initial positions ipos (3*np, scaled) in the
 * initial box init_box: no force nor energy when s > PairParams(3), otherwise
 * f = C s Vi^2 along the pair (C = PairParams(2), Vi = PairParams(4)) and e = f s r0 / 2.
 * Spring pairs have no cutoff test: every listed pair interacts. */
void ndo_compute23(int asym, int np, const double *pos, const int *typ, const double box[3],
                   const int *vnl, int neimax, int ntypes, const int *pairmat,
                   const double *P, const double *F, const double *rc2, const double *dr2,
                   double ptrmin, int ptsize, const int *spring, const double *pairparams,
                   const double *ipos, const double init_box[3], int cv, int cu,
                   double *force, double *pe, double virial[6], double *pene);

/* integrate1 (src/nd.f90:984-1046): drift, scaled displacement accumulation. */
void ndo_integrate1(int np, double *pos, const double *vel, const double *acc,
                    double *disp, const double box[3], double dt);

/* integrate2 (src/nd.f90:1081-1121): kick, acceleration, kinetic energy when ck. */
void ndo_integrate2(int np, double *vel, double *acc, const double *force, const int *typ,
                    const double *mass, double dt, int ck, double *ke);

/* mkIsoKE (src/nd.f90:1122-1158); returns the velocity scale applied. */
double ndo_mk_iso_ke(int np, double *vel, double dt, double keparm, double ketarget,
                     double kinetic);
/* mkIsoE (src/nd.f90:1159-1182); returns the velocity scale applied. */
double ndo_mk_iso_e(int np, double *vel, double dt, double eparm, double e0, double pe,
                    double ke);
/* mkIsoP (src/nd.f90:1183-1213): sequential box update from the virial. */
void ndo_mk_iso_p(double box[3], const double virial[6], double dt, double pparm,
                  const double ptarget[3]);

/* CheckListUpdate criterion (src/nd.f90:1447-1499).  Returns 1 to rebuild, 0 not,
 * -NDO_ERR_OUTSIDE for the reference's 'Particle outside of box'.  In SKIN mode the
 * displacement is converted to A with box and the box-strain term uses box_build. */
int ndo_check_list_update(int np, const double *disp, double skin, const double box[3],
                          const double box_build[3], double rc_max, int mode);

/* Run parameters for ndo_run: one 'Run n dt' of the reference deck (src/nd.f90:60-199). */
typedef struct {
  int nsteps;
  double dt;
  int out_steps;      /* OutFreq */
  int step0;          /* value of 'step' before the Run (StepNumber / previous Runs) */
  double skin;        /* < 0: 0.2*rcut as the reference */
  int neimax;         /* 0: reference estimate */
  int renei, renei_freq;
  int trigger;        /* NDO_TRIGGER_* */
  int brute;          /* force the O(N^2) list walk */
  int isokinetic; double ke_target, ke_tau; int ke_freq;
  int isobaric; double p_target[3], p_tau; int p_freq;
  int isoenergetic; double e_tau; int e_freq;
  int box_load; double box_vel[3];
  /* spring pairs (compute2/3): bit p set when pair id p+1 is a 'spring'; ipos (3*np) and
   * init_box are the initial positions and box of initPositionSet (src/nd.f90:1731-1763),
   * filled from pos/box at the start of the Run when ipos_set is 0, as dynamics does
   * (src/nd.f90:103) */
  int springs;
  double *ipos;
  double init_box[3];
  int ipos_set;
  /* Hold/Disp boundary conditions (src/nd.f90:1048-1075): hold[np] = hold group of each
   * atom (0: none); per group g = 1..nhold: bc_disp[g-1] (bcTyp 'd'), bc_dir[3(g-1)+a],
   * bc_parm[3(g-1)+a] (scaled displacement per step).  nhold = 0: none. */
  int nhold;
  const int *hold, *bc_disp, *bc_dir;
  const double *bc_parm;
} ndo_run_params;

/* One thermo line (src/read_input.f90:433-446) plus the full virial and box. */
typedef struct {
  int step;
  double pe, ke, eerr, pressure, volume;
  double virial[6];
  double box[3];
} ndo_thermo;

/* The 'dynamics' loop (src/nd.f90:60-199) on host arrays.  State (pos, vel, acc, box)
 * is advanced in place; force/pene receive the last step's values.  Thermo records for
 * step 1 of the run and every out_steps are written to rec (up to max_rec).  Returns
 * NDO_OK or an error code; nupdates receives the rebuild count. */
int ndo_run(const ndo_run_params *rp, int np, int ntypes, int npairs, int ptsize,
            double ptrmin, const double *mass, const int *pairmat, const double *pairparams,
            const double *P, const double *F, const double *rc2, const double *dr2,
            const int *typ, double *pos, double *vel, double *acc, double box[3],
            double *force, double *pene, ndo_thermo *rec, int max_rec, int *nrec,
            int *nupdates);

#ifdef __cplusplus
}
#endif
#endif
