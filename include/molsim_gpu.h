/*
 * molsim_gpu.h -- C-ABI of the MI355X (gfx950) non-bonded energy/force path for MOLSIM.
 *
 * The entry points mirror MOLSIM's own call surface for the charge / monoatomic branch:
 *
 *   msg_build_table  <- PotTwoBodyTab1/2/3        Src/potential.F90:868-1131
 *   msg_set_table    <- the ubuf/iubuflow/r2umin/r2atat state those routines leave behind
 *                       (Src/mol.F90:1087-1135, Src/particle.F90:2300-2308)
 *   msg_ewald_setup  <- EwaldSetup                Src/energy.F90:6601-7126
 *   msg_set_box      <- SetBoxParam + EwaldSetup + InitCellList after a box change (VolChange)
 *                                                 Src/mc.F90:3423-3436, Src/md.F90:699-700
 *   msg_nlist        <- NList / InitCellList / SetCellList   Src/nlist.F90:553-669,
 *                       Src/celllist.F90:76-410
 *   msg_utotal       <- UTotal(iStage)            Src/energy.F90:280-434
 *   msg_dutotal      <- DUTotal(lhsoverlap, lhepoverlap) for single-particle moves
 *                                                 Src/denergy.F90:98-279,1240-1543
 *   msg_commit       <- MCUpdate: EwaldUpdateArray + UpdateCellIp + ro <- rotm
 *                                                 Src/mc.F90:5203-5235, Src/denergy.F90:1776-1819,
 *                                                 Src/celllist.F90:366-382
 *   msg_set_weak_charges, msg_dutotal_charge, msg_commit_charge
 *                    <- lweakcharge: UWeakChargeA, ChargeChange's DUTotal and update
 *                                                 Src/energy.F90:7180-7276, Src/mc.F90:3683-3800,
 *                                                 Src/denergy.F90:154-164,1552-1699,3911-4145
 *
 * Conventions (those of the Fortran caller, so an iso_c_binding interface passes its
 * arrays straight through):
 *   - positions ro[3*i+d] (Fortran ro(3,np), real(8)), charges az[i], forces force[3*i+d];
 *   - particle types iptpn[i] and particle indices (ipnptm) are 1-based;
 *   - iubuflow[tp] is the 1-based start of type pair tp's records in ubuf (12 doubles each);
 *   - type pairs are numbered as MOLSIM's iptpt: (1,1),(1,2),...,(1,npt),(2,2),...;
 *   - energies in the caller's energy unit (kJ/mol), lengths in Angstrom; u_twob[0] is the
 *     sum over type pairs and u_twob[1..nptpt] the per-pair sums, as u%twob(0:nptpt);
 *     u_rec already contains EpsiFourPi, the self and the surface term, as u%rec.
 * Every function returns MSG_OK (0) or a negative MSG_ERR_* code; msg_last_error() gives text.
 * A context is used from one host thread at a time.
 */
#ifndef MOLSIM_GPU_H
#define MOLSIM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define MSG_MAX_TYPES 8

enum msg_status {
  MSG_OK = 0,
  MSG_ERR_ARG = -1,        /* bad argument / state                                     */
  MSG_ERR_R2UMIN = -2,     /* r2 < r2umin: below the table       (Src/energy.F90:710)  */
  MSG_ERR_IBUF = -3,       /* table scan past its end             (Src/energy.F90:717)  */
  MSG_ERR_TABLE = -4,      /* table construction did not converge (Src/potential.F90:1117) */
  MSG_ERR_TABLE_MEM = -5,  /* table larger than the buffer given                         */
  MSG_ERR_ODD_MESH = -6,   /* nmesh must be even                  (Src/energy.F90:6891) */
  MSG_ERR_MESH = -7,       /* nmesh below the order (or 4), above 1024, or too small a mesh
                              for a high order's spread workspace                          */
  MSG_ERR_HIP = -8,        /* HIP runtime error                                          */
  MSG_ERR_NOMEM = -9,      /* device allocation failed                                   */
  MSG_ERR_CELL = -10,      /* cell list: rcell <= 0 or position outside the box          */
  MSG_ERR_NLIST = -11,     /* Verlet list overflow (nneighpn > maxnneigh, Src/nlist.F90:752) */
  MSG_ERR_STATE = -12      /* call order: e.g. utotal before set_table                   */
};

enum msg_list_mode { MSG_LIST_ALL = 0, MSG_LIST_CELL = 1, MSG_LIST_VERLET = 2 };
enum msg_ewald_mode { MSG_EWALD_NONE = 0, MSG_EWALD_STD = 1, MSG_EWALD_SPM = 2 };
enum msg_precision {
  MSG_PREC_MIXED = 0,  /* FP32 pair/spread/FFT arithmetic, FP64 accumulators and Ewald terms */
  MSG_PREC_FP64 = 1    /* the same kernels in double: the plain FP64 GPU path               */
};

/* nmlPotential / nmlIntList / nmlSystem parameters the path needs (Src/potential.F90:307-392,
 * Src/nlist.F90:198-236). */
typedef struct msg_params {
  int np;              /* number of particles (one atom each)                  */
  int npt;             /* number of particle types, <= MSG_MAX_TYPES            */
  double boxlen[3];    /* box lengths (box-like cell, periodic in x, y, z)      */
  double rcut;         /* real-space cut-off                                    */
  double drnlist;      /* list skin; cell size >= rcut + drnlist                */
  double epsi4pi;      /* EpsiFourPi = Epsi0FourPi / relpermitt                 */
  int list_mode;       /* msg_list_mode (txintlist: vlist-equivalent / celllist / verlet) */
  int ewald_mode;      /* msg_ewald_mode (lewald, txewaldrec)                   */
  int lsurf;           /* surface (dipole) term                                 */
  double alpha;        /* ualpha                                                */
  int ncut;            /* std Ewald k-space cut-off                             */
  int ncutregion;      /* 0 sphere, 1 cube                                      */
  int nmesh;           /* SPME mesh points per dimension: any even number (the reference's
                          EwaldSetup, Src/energy.F90:6891-6892), order..1024        */
  int order;           /* SPME B-spline order, 2..11                            */
  int precision;       /* msg_precision                                         */
  double force_scale;  /* forces are multiplied by this at the end; UTotal uses
                          (sclene/scllen)/sclfor (Src/energy.F90:408); 1 keeps energy/length */
} msg_params;

/* Potential description for msg_build_table (form 0: 'default' with the 1/r term only,
 * Src/potential.F90:1266-1341; form 1: '(1,6,12)', Src/potential.F90:1349-1425). */
typedef struct msg_pot {
  int npt, form, lewald, pad;
  double alpha, rcut, epsi4pi;
  double utoltab, ftoltab, umaxtab, fmaxtab, r2uminin;
  double z[8], sig[8], eps[8], rad[8];
} msg_pot;

typedef struct msg_ctx msg_ctx;

const char* msg_last_error(void);

/* Host-side table construction (PotTwoBodyTab1/2/3 + SetUBuffer + CheckUBuffer,
 * Src/potential.F90:868-1258).  Outputs use the C-ABI conventions above (iubuflow 1-based).
 * Returns the number of doubles written to ubuf (> 0) or an error code. */
int msg_build_table(const msg_pot* pot, double* ubuf, int maxbuf, int* iubuflow, int* nugrid,
                    double* r2umin, double* r2atat);

int msg_create(const msg_params* params, msg_ctx** out);
void msg_destroy(msg_ctx* ctx);

/* Table state as left by PotTwoBodyTab (nptpt = npt*(npt+1)/2 entries per array). */
int msg_set_table(msg_ctx* ctx, const double* ubuf, int nbuf, const int* iubuflow,
                  const int* nugrid, const double* r2umin, const double* r2atat);

/* Charges (az) and 1-based particle types (iptpn).  Types are fixed for the context's life
 * time; charges change only through the weak-charge entry points below. */
int msg_set_particles(msg_ctx* ctx, const double* az, const int* iptpn);

/* Positions ro(3,np) from host memory, or from device memory already on this GPU. */
int msg_set_positions(msg_ctx* ctx, const double* ro);
int msg_set_positions_device(msg_ctx* ctx, const double* d_ro);

/* EwaldSetup: std k-factors or SPME influence function for the current box/parameters. */
int msg_ewald_setup(msg_ctx* ctx);

/* Box change, as SetBoxParam + EwaldSetup + InitCellList do in VolChange (Src/mc.F90:3423-3436)
 * and after a box rescaling in MD (Src/md.F90:699-700): boxlen[3] replaces the box, and the
 * cell grid, the Verlet list width and, after msg_ewald_setup, the Ewald set-up follow it.
 * The caller then sets the (scaled) positions.  Batches created before the change refuse to
 * run (MSG_ERR_STATE); execution units pick the new box up at their next launch. */
int msg_set_box(msg_ctx* ctx, const double* boxlen);

/* NList(iWriteInput/iSimulationStep): build the cell list (and the Verlet list in
 * MSG_LIST_VERLET mode) from the current positions. msg_utotal rebuilds the cell list
 * itself on every call; in Verlet mode it rebuilds the list on skin violation. */
int msg_nlist(msg_ctx* ctx);

/* Cell list after msg_nlist / msg_utotal: nc[3] cells per dimension, cell_start[ncell+1]
 * offsets (0-based) into cell_ip[np], the 1-based particle indices of each cell (cell id
 * ix + nc0*(iy + nc1*iz)) in the order MOLSIM's linked list visits them (iphead, ipnext, ...). */
int msg_get_cell_list(msg_ctx* ctx, int* nc, int* cell_start, int* cell_ip);

/* UTotal: u_twob[nptpt+1], *u_rec, *u_tot, force[3*np], *virial (any output may be NULL). */
int msg_utotal(msg_ctx* ctx, double* u_twob, double* u_rec, double* u_tot, double* force,
               double* virial);

/* Asynchronous UTotal on the context's stream; results stay on the device until
 * msg_fetch.  d_force (may be NULL) receives the forces in device memory. */
int msg_utotal_async(msg_ctx* ctx, double* d_force);
int msg_fetch(msg_ctx* ctx, double* u_twob, double* u_rec, double* u_tot, double* virial);

/* DUTotal for nptm moved particles ipnptm[0..nptm) (1-based) to rotm(3,nptm).
 * *lhsoverlap = 1 on a hard-core overlap (then only the flag is meaningful, as in the
 * reference).  du_twob[nptpt+1], *du_rec (x EpsiFourPi, with the surface term), *du_tot. */
int msg_dutotal(msg_ctx* ctx, int nptm, const int* ipnptm, const double* rotm, int* lhsoverlap,
                double* du_twob, double* du_rec, double* du_tot);

/* Batched DUTotal: ntrial independent single-particle trials against the SAME current
 * state (trial t moves ip[t] (1-based) to rotm(3,t)).  Outputs per trial. */
int msg_dutotal_batch(msg_ctx* ctx, int ntrial, const int* ip, const double* rotm,
                      int* lhsoverlap, double* du_tot, double* du_rec);

/* msg_dutotal_batch without allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  The trial
 * inputs and per-trial records go through a fixed device scratch array of the library (6.0 MB
 * per GPU, held until the process ends) instead of context buffers, so ntrial is at most
 * 16384 (MSG_ERR_ARG beyond).  Same results as msg_dutotal_batch bit for bit.  It needs the
 * state the allocating entry points set up: with SPME, one msg_dutotal / msg_dutotal_batch
 * call must have run on the context (else MSG_ERR_STATE).  All calls on one GPU share the
 * array: calls that overlap on one GPU (several host threads) race. */
int msg_dutotal_batch_noalloc(msg_ctx* ctx, int ntrial, const int* ip, const double* rotm,
                              int* lhsoverlap, double* du_tot, double* du_rec);

/* MCUpdate after an accepted move: positions, resident structure factors / mesh, cell list. */
int msg_commit(msg_ctx* ctx, int nptm, const int* ipnptm, const double* rotm);

/* Weak (titrating) charges, lweakcharge (UWeakChargeA Src/energy.F90:7180-7276; DUTotal's
 * weak-charge branch Src/denergy.F90:154-164,1552-1699,3911-4145): laz[i] (Fortran logical
 * as int) = 1 when particle i carries the charge zat[type] of its type (zat[npt]); its az
 * becomes zat[type] or 0, as IOCnf sets it (Src/molsim.F90:895-898).  A pair with an
 * uncharged particle has no pair energy; hard-core checks of trial moves still apply.
 * Std Ewald or no Ewald (the reference stops with SPME, Src/denergy.F90:1565-1570). */
int msg_set_weak_charges(msg_ctx* ctx, const int* laz, const double* zat);

/* ChargeChange trial (Src/mc.F90:3683-3800): the charge states of nptm (1 or 2) particles
 * ipnptm (1-based) flip, positions stay; outputs as msg_dutotal. */
int msg_dutotal_charge(msg_ctx* ctx, int nptm, const int* ipnptm, int* lhsoverlap,
                       double* du_twob, double* du_rec, double* du_tot);

/* MCUpdate of an accepted ChargeChange: laz and az of those particles flip
 * (Src/mc.F90:3781-3785). */
int msg_commit_charge(msg_ctx* ctx, int nptm, const int* ipnptm);

/* Velocity-Verlet MD on the device (VelVer, Src/md.F90:252-410): msg_md_init takes the
 * mass of each particle type (in the units that make force * force_scale / mass an
 * acceleration, as MOLSIM's (sclene/scllen)/sclfor scaling does) and the velocities
 * vel(3,np) (NULL = zero), and computes the initial accelerations (building the Verlet list
 * for the current positions, NList(iWriteInput), if none is valid).  msg_md_run advances
 * nsteps steps of length dt, each VelVer (drift, UTotal, kick) followed in MSG_LIST_VERLET
 * mode by NList(iSimulationStep) (Src/nlist.F90:597-621): with inlist = 0 the displacements
 * summed since the last build (drosum) decide, rebuilding when sqrt(r2max1) + sqrt(r2max2)
 * > drnlist with r2max1/r2max2 formed in the reference's particle order; with inlist > 0 the
 * list is rebuilt every inlist-th step of the macrostep.  A rebuild uses the positions of the
 * step just done.  log (may be NULL) receives [u_tot, kinetic energy] per step.  An error flag
 * raised by any step (list overflow, r2 < r2umin, position outside the box, table scan past
 * its end) is returned at the end of the call. */
int msg_md_init(msg_ctx* ctx, const double* mass, const double* vel);
int msg_md_run(msg_ctx* ctx, int nsteps, double dt, double* log);
int msg_get_positions(msg_ctx* ctx, double* ro);
int msg_get_velocities(msg_ctx* ctx, double* vel);
/* Number of Verlet-list builds so far (NList statistics, Src/nlist.F90:945-1074). */
long long msg_nlist_builds(msg_ctx* ctx);
/* inlist of nmlIntList (Src/nlist.F90:88-95,615-618): 0 (default) rebuilds by the skin test,
 * > 0 every inlist-th MD step of the macrostep. */
int msg_nlist_interval(msg_ctx* ctx, int inlist);
/* maxnneigh (Src/nlist.F90:240-246), the longest list a particle may have; a build that needs
 * more returns MSG_ERR_NLIST, as the reference stops (:752, :875-877).  0 (default) sizes it
 * from the density with a margin; the reference's max(np-1, ...) would take O(np^2) memory. */
int msg_set_nlist_width(msg_ctx* ctx, int maxnneigh);
/* NList(iBeforeMacrostep) (Src/nlist.F90:593-595): the displacement sums start again (without
 * a rebuild) and the macrostep's step count restarts. */
int msg_nlist_macrostep(msg_ctx* ctx);
/* The current Verlet list in MOLSIM's layout (built first for the current positions if none is
 * valid): kind 0 as SetVListMD (Src/nlist.F90:724-759, each pair once, kept by the particle the
 * (ip + jp) parity rule names), kind 1 as SetVListMC (:820-884, the full list).
 * jpnlist[(ip-1)*maxnneigh + k] (Fortran jpnlist(maxnneigh, np)) receives the 1-based
 * partners of particle ip in ascending order, nneighpn[ip-1] their number.  A list longer
 * than maxnneigh returns MSG_ERR_NLIST (the reference stops, :752). */
int msg_get_verlet_list(msg_ctx* ctx, int kind, int maxnneigh, int* nneighpn, int* jpnlist);

/* MOLSIM's random number generator (Random, Src/mollib.F90:2689-2713): *iseed as the
 * reference's iseed (negative or zero re-seeds), state[2] its two sequences (start with
 * {-1, -1}), *am its scale. */
double msg_random(int* iseed, int* state, double* am);

/* ntrial single-particle MC trials (MCPass / SPartMove / Metropolis / MCUpdate,
 * Src/mc.F90:1470-1750,5135-5235): particle chosen sequentially (isamp 0) or at random
 * (isamp 1); trial displacement (Random - 1/2) * dtran[type] per axis; Metropolis with
 * beta * dE (expmax = 87); accepted moves are committed.  The generator state (iseed,
 * rng_state, am) is read and updated.  Outputs: *du_sum (sum of accepted dE), *naccept,
 * events[t] (1 accepted, 0 rejected, -1 hard-core overlap; may be NULL).
 * Cell-list and all-pairs systems without weak charges run the pass on the device: one
 * workgroup, the generator on the device, an accepted move committed inside the launch
 * (UpdateCellIp on the cell list, S(k) for standard Ewald).  With SPME the launch corrects
 * phi for the moves it has accepted and ends after 16 of them; a UTotal pass recomputes phi
 * and the pass goes on (one host synchronisation per 16 accepted moves).  Verlet-list and
 * weak-charge systems run the trials from the host. */
int msg_mc_pass(msg_ctx* ctx, int ntrial, int isamp, const double* dtran, double beta,
                int* iseed, int* rng_state, double* am, double* du_sum, int* naccept,
                int* events);

/* Batched ensemble (no counterpart in the reference, which runs one system per process).
 * n resident cell-list systems (SPME or no Ewald) evaluated together: every stage of the
 * UTotal pass is one launch over the workgroups of all systems, captured once into a
 * hipGraph, so msg_batch_utotal is a single graph launch.  After msg_batch_sync,
 * msg_batch_results gives every system's energies in one download (u_tot, u_rec, virial:
 * arrays of n, any may be null) and msg_fetch / msg_get_forces a single system's.
 * msg_batch_dutotal runs ntrial[s] independent trials of every system s (indices 1-based,
 * indices and positions concatenated over systems in batch order) as one launch with one
 * synchronisation, after a batched pass if any system moved.
 * Contract: a member context used on its own must be synchronised (msg_fetch, msg_utotal)
 * before the next batched call, and the batch synchronised (msg_batch_sync or a call that
 * returns results) before a member is used on its own.  Tiles (msg_set_tile) and Ewald
 * set-up are read at msg_batch_create. */
typedef struct msg_batch msg_batch;
int msg_batch_create(msg_ctx** ctxs, int n, msg_batch** out);
void msg_batch_destroy(msg_batch* batch);
int msg_batch_utotal(msg_batch* batch);
int msg_batch_sync(msg_batch* batch);
int msg_batch_query(msg_batch* batch);
/* One MCPass per system of a batch, the chains side by side on the device (one 1024-thread
 * workgroup each, one launch per SPME order present); with SPME, one batched UTotal pass
 * after every launch in which a chain filled its log of accepted moves (32 at orders up to
 * 6, else 16) (no counterpart
 * in the reference, which runs one chain per process).  Every system must be SPME or
 * Ewald-free without weak charges.  Per system s: ntrial trials as msg_mc_pass;
 * dtran[s * MSG_MAX_TYPES + t] its displacement per type, iseed[s], rng_state[2 s .. 2 s + 1]
 * and am[s] its generator state (read and updated), du_sum[s] and naccept[s] its results,
 * events[s * ntrial + t] its decisions (each may be NULL).  After a batched refresh the cells
 * are sorted afresh (SetCellList order) rather than kept in UpdateCellIp order, so
 * real-space sums may round differently from a chain run alone. */
int msg_batch_mc_pass(msg_batch* batch, int ntrial, int isamp, const double* dtran, double beta,
                      int* iseed, int* rng_state, double* am, double* du_sum, int* naccept,
                      int* events);
int msg_batch_results(msg_batch* batch, double* u_tot, double* u_rec, double* virial);
int msg_batch_dutotal(msg_batch* batch, const int* ntrial, const int* ip, const double* rotm,
                      int* lhsoverlap, double* du_tot);

/* Execution unit for continuous batching (no counterpart in the reference): one batch of
 * requests at a time on the unit's own streams -- a UTotal pass over ut[0..n_u) (energies;
 * forces downloaded when want_forces) and ntrial[s] MC trials of each dec[s] (ip 1-based
 * and rotm concatenated over dec), as segmented launches over all its systems.  Launch
 * returns at once; several units in flight overlap on the device.  A system must not be in
 * two units in flight at once.  msg_unit_results waits and copies: u_tot/u_rec/virial
 * (n_u each), force (concatenated force(3,np) of the ut systems), lhsoverlap/du_tot (all
 * trials in launch order); any may be null. */
typedef struct msg_unit msg_unit;
int msg_unit_create(msg_unit** out);
void msg_unit_destroy(msg_unit* unit);
int msg_unit_launch(msg_unit* unit, msg_ctx** ut, int n_u, int want_forces, msg_ctx** dec, int n_d,
                    const int* ntrial, const int* ip, const double* rotm);
int msg_unit_query(msg_unit* unit);
int msg_unit_results(msg_unit* unit, double* u_tot, double* u_rec, double* virial, double* force,
                     int* lhsoverlap, double* du_tot);

/* 1 when everything launched on the context's stream has completed, 0 while work is pending. */
int msg_query(msg_ctx* ctx);

/* Pair-kernel tile of a cell-list system (no counterpart in the reference): pair_lanes lanes
 * per home particle, 256 / pair_lanes homes per workgroup (4, 8, 16, 32 or 64; the default
 * follows the expected candidates per home), 0 for the wave-per-cell-chunk kernel, or 108 /
 * 116 for the cell-tile kernel (mixed precision; 8 or 16 lanes per home, a tile's neighbour
 * cells staged in LDS).  Results do not depend on the tile beyond FP32 summation order.  Set
 * by the autotuner per size bucket; systems of more than 4 particle types ignore it. */
int msg_set_tile(msg_ctx* ctx, int pair_lanes);
int msg_get_tile(msg_ctx* ctx);

/* Forces force(3,np) of the last UTotal pass (scaled like msg_utotal's), e.g. after a batch. */
int msg_get_forces(msg_ctx* ctx, double* force);

/* Stream the context launches on (a hipStream_t), for callers that time kernels. */
void* msg_stream(msg_ctx* ctx);

/* Per-kernel timing with HIP events on the context's stream.  Kernel ids: */
enum msg_kernel_id {
  MSG_K_BIN = 0, MSG_K_SCAN, MSG_K_PAIR, MSG_K_SPREAD, MSG_K_FFT_X, MSG_K_FFT_Y,
  MSG_K_FFT_ZCONV, MSG_K_FFT_YI, MSG_K_FFT_XI, MSG_K_GATHER, MSG_K_REDUCE, MSG_K_EWSTD,
  MSG_K_DE, MSG_K_VLIST, MSG_K_SCATTER, MSG_K_FINALIZE, MSG_K_COUNT
};
int msg_profile(msg_ctx* ctx, int enable);
int msg_profile_read(msg_ctx* ctx, int kernel_id, double* total_ms, int* launches);
const char* msg_kernel_name(int kernel_id);

/* Bytes of device memory the context holds (for admission by HBM footprint). */
long long msg_device_bytes(msg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
