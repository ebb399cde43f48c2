/*
 * phx_elph.h - C ABI of the MI355X (gfx950) electron-phonon linewidth path.
 *
 * Drop-in replacement for Phoebe's el-ph scattering hot path:
 *   InteractionElPhWan (src/interaction/interaction_elph.h:30-219),
 *   kokkosZHEEV (src/common_kokkos.h:49) and the batched Bloch
 *   diagonalizations (src/harmonic/harmonic.h:80-91), plus the linewidth
 *   builder ElScatteringMatrix::builder, switchCase 2
 *   (src/bte/el_scattering.cpp:38-473).
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles;
 *   - matrices are column-major (Eigen's default), one matrix contiguous;
 *   - energies in Ry, lengths in Bohr, wavevectors Cartesian incl. 2*pi;
 *   - every function returns 0 on success and a negative status on error;
 *     phx_last_error() returns the message of the calling thread's last error
 *     (the reference prints and exits through Error(), exceptions.cpp:6-16;
 *     the adaptor in phoebe_amd/csrc/phx_adaptor.hpp maps status != 0 back
 *     to that behaviour).
 * All pointers are host pointers unless a name ends in _dev.
 */
#ifndef PHX_ELPH_H
#define PHX_ELPH_H

#ifdef __cplusplus
extern "C" {
#endif

/* precision of the Fourier-transform and rotation GEMMs; phases,
 * eigensolves and the linewidth accumulation are always FP64.  PHX_BF16X3
 * keeps FP32 data and runs the last mesh-transform GEMMs as split BF16
 * (hi * hi + hi * lo + lo * hi, FP32 accumulation); linewidth builder with the
 * separable transform only. */
enum { PHX_FP64 = 0, PHX_FP32 = 1, PHX_BF16X3 = 2 };

/* how the R_ph sum of calcCouplingSquared is evaluated in the linewidth builder:
 * DIRECT    - one GEMM over the R_ph vectors for every (k1, k2) pair (the
 *             reference's algorithm, interaction_elph.cpp:357-378);
 * SEPARABLE - k2 on the Gamma-centred mesh: e^{ik2.R} = prod_d w_d^{j_d n_d},
 *             three short lattice-direction transforms over the whole mesh
 *             (exact, same FP64 accuracy, far fewer operations). */
enum { PHX_TRANSFORM_DIRECT = 0, PHX_TRANSFORM_SEPARABLE = 1 };

/* where band velocities are set to zero as "Gamma": the reference's batched
 * GPU forms use |k|^2 <= 1e-6, its single-point CPU forms |k| < 1e-6. */
enum { PHX_GAMMA_BATCHED = 0, PHX_GAMMA_POINT = 1 };

/* Message of the last error on the calling thread ("" if none). */
const char* phx_last_error(void);

/* Library version string. */
const char* phx_version(void);

/* Number of HIP devices this process sees (0 when no GPU or driver is
 * present: *count = 0 and the call still returns 0).  No counterpart in the
 * reference, which gets its device from Kokkos::initialize. */
int phx_device_count(int* count);

/* ---------------------------------------------------------------- eigensolver
 * Batched Hermitian eigen-decomposition, in place.
 * Replaces kokkosZHEEV(StridedComplexView3D& A, DoubleView2D& W)
 * (src/common_kokkos.h:49, common_kokkos.cpp:11-100).
 *   A  : M matrices N x N, column-major complex; on return the eigenvectors
 *        (columns), eigenvalues ascending in W (M x N).
 *   info (optional, M ints): Jacobi sweeps used, -1 if not converged.
 * N <= 44 (the matrix and its eigenvectors stay in one wavefront's LDS). */
int phx_zheev_batched(int M, int N, double* A, double* W, int* info);

/* Same on device buffers, enqueued on a HIP stream (0 = default stream). */
int phx_zheev_batched_dev(int M, int N, double* A_dev, double* W_dev, int* info_dev,
                          void* stream);

/* ---------------------------------------------------------------- Bloch FTs
 * ElectronH0Wannier::kokkosBatchedDiagonalizeFromCoordinates
 * (src/harmonic/electron_h0_wannier.cpp:777-877):
 *   H(k) = sum_R e^{ik.R}/deg(R) h0R(R,:,:), then eigen-decomposition.
 *   k    : nk x 3 Cartesian;  R: nR x 3 Cartesian;  deg: nR;
 *   h0R  : (nR, nw, nw) complex, row-major (h0R(R,m,n) of the reference);
 *   energies: nk x nw;  eigvecs: nk x (nw x nw column-major complex). */
int phx_el_bloch_diag(int nk, const double* k, int nw, int nR, const double* R,
                      const double* deg, const double* h0R, double* energies,
                      double* eigvecs);

/* phx_el_bloch_diag without any allocation: the same kernels and results,
 * with every device array (the copies of k, R, deg and h0R, the Bloch
 * matrices, eigenvalues and convergence flags) in one fixed 64 MiB device
 * array per GPU, held until the process ends, instead of hipMalloc'd
 * buffers; no hipMalloc, hipFree or host heap allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Needs
 * 24 nk + 32 nR + 16 nw^2 (nk + nR) + 8 nk nw + 4 nk bytes, each array rounded
 * up to 256 (nw = 16: about 15000 points per call with nR = 1000); larger
 * calls return an error.  Calls that overlap on one GPU share the array and
 * race. */
int phx_el_bloch_diag_noalloc(int nk, const double* k, int nw, int nR, const double* R,
                              const double* deg, const double* h0R, double* energies,
                              double* eigvecs);

/* PhononH0::kokkosBatchedDiagonalizeFromCoordinates
 * (src/harmonic/phonon_h0_kokkos.cpp:6-219), no long-range term:
 *   D(q) = sum_R e^{-iq.R} w(R) mat2R(R,:,:), hermitized, divided by
 *   sqrt(M_m M_n); omega = sign(l) sqrt|l|; eigenvectors divided by sqrt(M_m)
 *   when mass_scaling != 0.
 *   mat2R: (nR, nm, nm) real row-major;  masses: nm (Ry units);
 *   omega: nq x nm;  eigvecs: nq x (nm x nm column-major complex). */
int phx_ph_bloch_diag(int nq, const double* q, int nm, int nR, const double* R,
                      const double* w, const double* mat2R, const double* masses,
                      int mass_scaling, double* omega, double* eigvecs);

/* Timing form of phx_el_bloch_diag (phonon = 0) and phx_ph_bloch_diag
 * (phonon != 0, masses given, mass scaling on): inputs uploaded once, then
 * reps repetitions of the Bloch build and the batched eigensolve on device
 * buffers, each kernel bracketed by HIP events.  ms[0], ms[1]: average
 * milliseconds per launch of the Bloch build and of the eigensolve; the
 * results of the last repetition are returned as the untimed calls do.  No
 * counterpart in the reference (a measurement entry point). */
int phx_bloch_diag_timed(int phonon, int nk, const double* k, int n, int nR, const double* R,
                         const double* w, const double* coef, const double* masses, int reps, double* ms,
                         double* energies, double* eigvecs);

/* ---------------------------------------------------------------- InteractionElPhWan */
typedef struct phx_elph phx_elph;

/* InteractionElPhWan(Crystal&, const Eigen::Tensor<cd,5>& gW, elBravaisVectors,
 * elDegeneracies, phBravaisVectors, phDegeneracies, PhononH0*)
 * (interaction_elph.h:93-100, interaction_elph.cpp:10-62).
 *   gW : the Eigen tensor (nw(w2), nw(w1), nm, nRp, nRe) column-major, i.e.
 *        (nRe, nRp, nm, w1, w2) row-major;  elR: nRe x 3;  phR: nRp x 3.
 * The polar (Froehlich) term is off until phx_elph_set_polar gives its
 * parameters (the reference's constructor turns it on for a dielectric tensor
 * and more than one species, :24-32; phx_polar_in_use states that rule). */
int phx_elph_create(phx_elph** h, const double* gW, int nw, int nm, int nRp, int nRe,
                    const double* elR, const double* elDeg, const double* phR,
                    const double* phDeg, int precision);

/* Same, with gW generated on the device from the synthetic recipe
 * gW = decay[Re,Rp] * (n(2f) + i n(2f+1)), n = counter-hash normal deviate
 * (phoebe_amd/synth.py), so that tens of GB never cross PCIe. */
int phx_elph_create_synthetic(phx_elph** h, int nw, int nm, int nRp, int nRe,
                              const double* elR, const double* elDeg, const double* phR,
                              const double* phDeg, const double* decay,
                              unsigned long long seed, int precision);

/* InteractionElPhWan::cacheElPh(const Eigen::MatrixXcd& eigvec1 (nw x nb1),
 *                               const Eigen::Vector3d& k1C)
 * (interaction_elph.h:161, interaction_elph.cpp:492-701). */
int phx_elph_cache(phx_elph* h, const double* U1, int nb1, const double* k1);

/* InteractionElPhWan::calcCouplingSquared(eigvec1, eigvecs2, eigvecs3, q3Cs,
 *                                         polarData)
 * (interaction_elph.h:149-154, interaction_elph.cpp:229-448).
 *   U2s  : concatenation of the nq matrices (nw x nb2s[i]) column-major;
 *   e3s  : nq x (nm x nm) column-major;  q3: nq x 3 Cartesian;
 *   polar: NULL (the reference's usePolarCorrection = false), or nq x nm
 *          complex: polarCorrectionPart1 of each q (phx_elph_polar_part1);
 *          gFinal(nu, b1, b2) += polar(q, nu) sum_w conj(U2(w, b2)) U1(w, b1)
 *          before |g|^2 for every q with |q3| > 1e-8 (:285-312, :406-413,
 *          polarCorrectionPart2 :205-227);
 *   out  : concatenation of the nq tensors (nb1, nb2s[i], nm) column-major,
 *          the getCouplingSquared(ik) layout (interaction_elph.cpp:434-446). */
int phx_elph_coupling_sq(phx_elph* h, int nq, const double* U2s, const int* nb2s,
                         const double* e3s, const double* q3, const double* polar,
                         double* out);

/* ---------------------------------------------------------------- polar term
 * The Froehlich long-range part of the coupling (Verdi and Giustino, PRL 115,
 * 176401, Eq. 4), as InteractionElPhWan's polarCorrectionPart1Static
 * (interaction_elph.cpp:152-203) computes it. */
typedef struct {
  int natoms;
  const double* atomPos;     /* natoms x 3, Cartesian Bohr */
  const double* bornCharges; /* natoms x 3 x 3: Z(atom, i, pol), gqDotZ = sum_i (q+G)_i Z(atom, i, pol) */
  const double* dielectric;  /* 3 x 3 */
  double recip[9];           /* reciprocal cell, column-major, columns b_i */
  double volume;             /* unit-cell volume, Bohr^3 */
  int coarseMesh[3];         /* the force constants' q grid: G = sum_i m_i b_i, |m_i| <= coarseMesh[i] */
} phx_polar;

/* Whether the reference turns the polar term on (interaction_elph.cpp:24-32):
 * a dielectric tensor with squared norm > 1e-10 and more than one species.
 * Returns 1 or 0. */
int phx_polar_in_use(const double* dielectric, int num_species);

/* Stores the polar parameters on the interaction (the Crystal and PhononH0
 * fields polarCorrectionPart1 reads, :140-151); NULL clears them. */
int phx_elph_set_polar(phx_elph* h, const phx_polar* polar);

/* InteractionElPhWan::polarCorrectionPart1(q3, ev3) for nq wavevectors on the
 * GPU: x(nu) = sum over q+G with 0 < (q+G).eps.(q+G) < 4 gMax (gMax = 14) of
 * i 2 fourPi / V e^{-(q+G).eps.(q+G)/4} / ((q+G).eps.(q+G))
 * e^{-i (q+G).tau_a} ((q+G).Z(a, :, pol)) e3(3 a + pol, nu), with
 * fourPi = 2 pi as constants.h:12 defines it.
 *   q3 : nq x 3 Cartesian;  e3s: nq x (nm x nm) column-major;
 *   out: nq x nm complex. */
int phx_elph_polar_part1(phx_elph* h, int nq, const double* q3, const double* e3s, double* out);

/* InteractionElPhWan::getCouplingDimensions() (interaction_elph.cpp:450-456):
 * the device extents (nRe, nRp, nm, nw, nw). */
int phx_elph_get_coupling_dims(phx_elph* h, int* dims5);

/* InteractionElPhWan::estimateNumBatches(nk2, nb1) (interaction_elph.cpp:458-490)
 * with the reference's memory model: ceil(nk2 * bytes per q / available).
 *   available_bytes > 0: the caller's DeviceManager::getAvailableMemory()
 *     (common_kokkos.cpp:131-133), i.e. MAXMEM minus what every object
 *     registered, this one included (phx_elph_device_memory_usage);
 *   available_bytes <= 0: a DeviceManager holding only this object, MAXMEM
 *     from the environment in GB, default 16 (common_kokkos.cpp:104-112).
 * Returns the number of batches (0 when nk2 = 0), or a negative status with
 * "Insufficient memory!" when one q does not fit (:477-488). */
int phx_elph_estimate_num_batches(phx_elph* h, int nk2, int nb1, double available_bytes);

/* InteractionElPhWan::getDeviceMemoryUsage() (interaction_elph.cpp:703-707):
 * the bytes the reference registers with its DeviceManager for this object
 * (16 per element of the coupling tensor and of elPhCached after a
 * cacheElPh, 8 per R-vector component and degeneracy). */
int phx_elph_device_memory_usage(phx_elph* h, double* bytes);

int phx_elph_destroy(phx_elph* h);

/* ---------------------------------------------------------------- harmonic models
 * The inputs of the band structures, read from the files Phoebe reads, with
 * the set-up work of the reference's constructors.  Host code, except the
 * *_diag calls, which run on the GPU. */
typedef struct phx_ph_model phx_ph_model;
typedef struct phx_el_model phx_el_model;

/* QEParser::parsePhHarmonic (qe_input_parser.cpp:385-562) + the PhononH0
 * constructor (phonon_h0.cpp:18-235): QE force constants (q2r.x), the acoustic
 * sum rule ("" none, "simple"; "crystal" is not supported), the Wigner-Seitz
 * reorder (:426-535, 672-771) and the long-range set-up (:53-160). */
int phx_ph_model_read_qe(phx_ph_model** m, const char* fc_path, const char* sum_rule);
/* dims8 = {natoms, nm, nR, nG, hasDielectric, qGrid[3]} */
int phx_ph_model_dims(phx_ph_model* m, int* dims8);
/* cell9 / recip9 column-major (columns a_i / b_i, Bohr / Bohr^-1), volume,
 * positions natoms x 3 (Bohr), masses nm (Ry mass units); any may be NULL. */
int phx_ph_model_crystal(phx_ph_model* m, double* cell9, double* recip9, double* volume,
                         double* positions, double* masses);
/* R nR x 3, weights nR, mat2R nR x nm x nm (R, iAt*3+i, jAt*3+j); any may be NULL. */
int phx_ph_model_force_constants(phx_ph_model* m, double* R, double* weights, double* mat2R);
/* born natoms x 9, dielectric 9, gVectors nG x 3, lrConstant natoms x 9, norm_gmax2 = {lrNorm, gMax} */
int phx_ph_model_long_range(phx_ph_model* m, double* born, double* dielectric, double* gVectors,
                            double* lrConstant, double* norm_gmax2);
/* PhononH0::kokkosBatchedDiagonalizeFromCoordinates (phonon_h0_kokkos.cpp:177-207)
 * with the long-range term: omega nq x nm, eigvecs nq x (nm x nm) column-major,
 * divided by sqrt(M) when mass_scaling. */
int phx_ph_model_diag(phx_ph_model* m, int nq, const double* q, int mass_scaling, double* omega,
                      double* eigvecs);
/* PhononH0::kokkosBatchedDiagonalizeWithVelocities (phonon_h0_kokkos.cpp:378-532):
 * as phx_ph_model_diag with mass scaling, plus the velocity operator
 * velocities nq x nm x nm x 3 complex, (q, m, n, i) row-major (the reference's
 * ComplexView4D): central differences (1e-8) of sqrt(D), rotated to the
 * eigenbasis, zero at Gamma, degenerate blocks (1e-4 cm^-1) diagonalized per
 * direction (harmonic.cpp:3-203; blocks of up to 8 modes).
 * gamma_rule: PHX_GAMMA_BATCHED (velocities zero unless |q|^2 > 1e-6, the
 * batched GPU form) or PHX_GAMMA_POINT (zero when |q| < 1e-6, the single-point
 * form PhononH0::diagonalizeVelocityFromCoordinates, phonon_h0.cpp:895-1070). */
int phx_ph_model_diag_velocities(phx_ph_model* m, int nq, const double* q, int gamma_rule,
                                 double* omega, double* eigvecs, double* velocities);
int phx_ph_model_destroy(phx_ph_model* m);

/* QEParser::parseElHarmonicWannier (qe_input_parser.cpp:883-1040): Wannier90
 * seedname_tb.dat.  cell9: the crystal's direct cell (column-major, Bohr), as
 * the reference passes the phonon crystal; NULL uses the file's cell. */
int phx_el_model_read_w90(phx_el_model** m, const char* tb_path, const double* cell9);
/* dims2 = {nw, nR} */
int phx_el_model_dims(phx_el_model* m, int* dims2);
/* cell9, R nR x 3 (Cartesian Bohr), deg nR, h0R nR x nw x nw complex (Ry),
 * rMatrix 3 x nR x nw x nw complex (Bohr); any may be NULL. */
int phx_el_model_arrays(phx_el_model* m, double* cell9, double* R, double* deg, double* h0R,
                        double* rMatrix);
/* ElectronH0Wannier::kokkosBatchedDiagonalizeFromCoordinates
 * (electron_h0_wannier.cpp:777-877): energies nk x nw, eigvecs nk x (nw x nw). */
int phx_el_model_diag(phx_el_model* m, int nk, const double* k, double* energies, double* eigvecs);
/* ElectronH0Wannier::kokkosBatchedDiagonalizeWithVelocities
 * (electron_h0_wannier.cpp:888-999): energies, eigenvectors and the velocity
 * operator nk x nw x nw x 3 complex, (k, m, n, i) row-major: central
 * differences (1e-8) of H(k), degenerate blocks (1 micro-eV) diagonalized per
 * direction (blocks of up to 8 bands).  gamma_rule as for phonons; the
 * single-point form is ElectronH0Wannier::batchedDiagonalizeWithVelocities
 * (electron_h0_wannier.cpp:633-760). */
int phx_el_model_diag_velocities(phx_el_model* m, int nk, const double* k, int gamma_rule,
                                 double* energies, double* eigvecs, double* velocities);
int phx_el_model_destroy(phx_el_model* m);

/* ---------------------------------------------------------------- el-ph file
 * Phoebe's prefix.phoebe.elph.hdf5, read and written without libhdf5 (the
 * library's own HDF5 subset: root-group datasets, contiguous or compact
 * storage, no filters).  Host code only; no GPU needed. */
typedef struct phx_elph_file phx_elph_file;

/* parseHDF5 + parseHeaderHDF5 (interaction_elph_parsing.cpp:626-670, 161-233):
 * opens the file, decides the format (a file without /fileFormat is format 1,
 * :651-663) and reads the header datasets.  A missing file fails with the
 * reference's message (:629-636). */
int phx_elph_file_open(phx_elph_file** f, const char* path);

/* ints13 = {fileFormat, numElectrons, numSpin, numElBands (nw), numPhModes (nm),
 *           nRe, nRp, kMesh[3], qMesh[3]}  (meshes 0 when the file has none). */
int phx_elph_file_header(phx_elph_file* f, int* ints13);

/* elBravaisVectors / elDegeneracies / phBravaisVectors / phDegeneracies
 * (interaction_elph_parsing.cpp:199-211): elR nRe x 3, phR nRp x 3 (Bohr);
 * any pointer may be NULL. */
int phx_elph_file_bravais(phx_elph_file* f, double* elR, double* elDeg, double* phR,
                          double* phDeg);

/* gW of the lattice vectors R_e in [re_begin, re_begin + re_count), in the
 * phx_elph_create layout: the slice a rank or pool selects
 * (format 1: interaction_elph_parsing.cpp:416-425, 495-502; format 2: :576-582). */
int phx_elph_file_read_gw(phx_elph_file* f, int re_begin, int re_count, double* out);

int phx_elph_file_close(phx_elph_file* f);

/* writeElPhCouplingHDF5v1 / v2 + writeHeaderHDF5 (qe_to_phoebe_utils.cpp:446-626,
 * 628-720, 377-443): format 1 (one /gWannier) or 2 (/gWannier_<irE>).
 *   ints12 = {numElectrons, numSpin, nw, nm, nRe, nRp, kMesh[3], qMesh[3]};
 *   elR nRe x 3, phR nRp x 3, gW in the phx_elph_create layout.
 * An existing file at path is replaced (qe_to_phoebe_utils.cpp:469-474). */
int phx_elph_file_write(const char* path, int format, const int* ints12, const double* elR,
                        const double* elDeg, const double* phR, const double* phDeg,
                        const double* gW);

/* InteractionElPhWan::parse (interaction_elph_parsing.cpp:672-692) followed by
 * the constructor: the coupling tensor is streamed from the file into device
 * memory through two pinned staging buffers (reads overlap the copies; FP32
 * handles convert each chunk on the device), so no host copy of gW is made. */
int phx_elph_create_from_file(phx_elph** h, const char* path, int precision);

/* ---------------------------------------------------------------- linewidths
 * Whole electron linewidth calculation of ElScatteringMatrix::builder,
 * switchCase 2 (src/bte/el_scattering.cpp:38-473), inner == outer band
 * structure on a Gamma-centred mesh with an energy window
 * (HelperElScattering storedAllQ3Case2, helper_el_scattering.cpp:58-162),
 * Gaussian smearing, no symmetries, no polar term. */
typedef struct {
  double recip[9]; /* reciprocal cell, column-major, columns b_i (2*pi*A^-T) */
  int mesh[3];     /* k mesh = q mesh */
  int nw, nRh;
  const double *Rh, *degh, *h0R; /* electron Wannier Hamiltonian */
  int nm, nRf;
  const double *Rf, *wf, *mat2R, *masses; /* reordered force constants */
  int nRe, nRp;
  const double *Re, *degE, *Rp, *degP;
  const double* gW;          /* host gW, or NULL to use the synthetic recipe: */
  const double* gwDecay;     /*   (nRe, nRp) */
  unsigned long long gwSeed;
  int nCalc;                 /* (T, mu) pairs, Ry */
  const double *temperatures, *chemPots;
  double sigma, winMin, winMax; /* Gaussian width, energy window (Ry) */
  /* rigid-ion long-range term of D(q), present when the force constants come
   * with a dielectric tensor (PhononH0, phonon_h0.cpp:53-160 and
   * phonon_h0_kokkos.cpp:48-149); nG = 0 leaves it out. */
  int natoms, nG;
  const double* atomPos;     /* natoms x 3, Cartesian Bohr */
  const double* bornCharges; /* natoms x 3 x 3, (atom, i, j), after the sum rule */
  const double* dielectric;  /* 3 x 3, (i, j) */
  const double* gVectors;    /* nG x 3 */
  const double* lrConstant;  /* natoms x 3 x 3, (atom, i, j): added to the diagonal blocks */
  double lrNorm, gMax;       /* e2 * fourPi / volume (fourPi as constants.h:12 defines it), 14 */
  /* smearing: PHX_SMEAR_GAUSSIAN (width sigma) or PHX_SMEAR_ADAPTIVE
   * (AdaptiveGaussianDeltaFunction, delta_function.cpp:55-97: per-pair width
   * from the group velocities of the two electron states, el_scattering.cpp:
   * 291-295; sigma is then unused and the energy screens are off). */
  int smearing;
} phx_problem;

enum { PHX_SMEAR_GAUSSIAN = 0, PHX_SMEAR_ADAPTIVE = 1 };

typedef struct phx_lw phx_lw;

/* Band structures, window, phonons on the q-set, device-resident operands. */
int phx_lw_setup(phx_lw** h, const phx_problem* prob, int precision, int k1_batch,
                 int transform);
/* Same, with the coupling (nRe, nRp, Re, degE, Rp, degP, gW) taken from Phoebe's
 * el-ph file instead of prob: the builder's interaction comes from
 * InteractionElPhWan::parse (interaction_elph_parsing.cpp:672-692).  The file's
 * numElBands / numPhModes must equal prob->nw / prob->nm; gW is streamed from the
 * file into device memory. */
int phx_lw_setup_file(phx_lw** h, const phx_problem* prob, const char* elph_path, int precision,
                      int k1_batch, int transform);

/* Points prob's band-structure fields (recip from the phonon crystal, the
 * electron Hamiltonian, the force constants and the long-range term) at the
 * models' arrays, which must outlive prob's use; the coupling, mesh, (T, mu),
 * smearing and window fields are left to the caller.  As in the reference's
 * apps (electron_wannier_transport_app.cpp:16-20), the electron model should
 * be read with the phonon crystal's cell. */
int phx_problem_set_models(phx_problem* prob, phx_el_model* el, phx_ph_model* ph);
int phx_lw_num_active(phx_lw* h);

int phx_lw_num_states(phx_lw* h);
/* active points (full-mesh indices), band extrema, energies of the states */
int phx_lw_get_window(phx_lw* h, int* active, int* band_lo, int* band_hi, double* energies);
/* phonon frequencies (nk x nm, full mesh order) and the q-set mask (nk bytes) */
int phx_lw_get_phonons(phx_lw* h, double* omega, unsigned char* in_set);
/* Accumulate the scattering of the outer points k1_list (active indices;
 * NULL = all) against inner points [k2_begin, k2_end) (k2_end < 0: all).
 * Enqueued on the handle's stream; returns the number of (k1,k2) pairs. */
long long phx_lw_run(phx_lw* h, int nk1, const int* k1_list, int k2_begin, int k2_end);
/* Linewidths (nCalc x numStates, Ry), with degeneracyAveragingLinewidths
 * (scattering.cpp:1168-1207) applied to row 0 when average != 0. */
int phx_lw_get_linewidths(phx_lw* h, double* out, int average);
int phx_lw_reset(phx_lw* h);
/* Energy-conservation screen of the separable transform: an inner point k2 is
 * skipped for a panel (k1, b1) when every band e2 of k2 and every mode omega
 * >= 5 cm^-1 at q = k2 - k1 give |e1 - e2 +- omega| >= nsigma * sigma, i.e.
 * when all its Gaussian weights (delta_function.cpp:40-45) are below
 * exp(-nsigma^2) of their peak; inside the rows it keeps, a (band, mode)
 * pair whose two weights are both that small is skipped the same way.  0
 * turns it off (the default); otherwise nsigma >= 6.  No counterpart in the
 * reference, which evaluates every pair. */
int phx_lw_set_screen(phx_lw* h, double nsigma);

/* The polar term of the linewidth builder (el_scattering.cpp through
 * helper_el_scattering.cpp:170-195 and calcCouplingSquared :285-312,
 * :406-413): polarCorrectionPart1 of every q of the q-set (folded mesh
 * wavevector, its mass-scaled eigenvectors) is computed on the GPU, and every
 * pair with k2 != k1 adds polar(q, nu) (U2^H U1)(b2, b1) to gFinal.  Runs with
 * the direct transform (the separable transform raises an error with it on);
 * NULL turns it off.  Call after phx_lw_setup. */
int phx_lw_set_polar(phx_lw* h, const phx_polar* polar);
/* ---- Sweeps without host round trips.  Prepare every batch of a sweep over
 * k1_list (all active points when null; same batching as phx_lw_run) with
 * its own device metadata and all scratch sized; returns the batch count.
 * No counterpart in the reference, whose k1 loop is serial on the host. */
long long phx_lw_plan(phx_lw* h, int nk1, const int* k1_list, int k2_begin, int k2_end);
/* Enqueue the prepared sweep on the handle's stream without host
 * synchronization; graph != 0 captures it into a hipGraph on first use and
 * replays that graph.  Handles run concurrently on their own streams.
 * Returns the (k1, k2) pair count. */
long long phx_lw_plan_run(phx_lw* h, int graph);
/* Batch shapes: outer points per GEMM-form batch (1..32, the cacheElPh GEMM
 * M tile is 2 x k1_batch) and mesh lines per line batch (1..16). */
int phx_lw_set_batching(phx_lw* h, int k1_batch, int lines_per_batch);
/* Autotuner: times full runs over the sample k1_list for k1 batches {8, 16,
 * 32}, lines per batch {1, 2, 4} (line batching on) and panels per separable
 * launch {16, 48, 128}, one knob at a time, keeps the fastest; the
 * linewidths are restored.  out4 = {k1_batch, lines_per_batch, panel_chunk,
 * best seconds}.  No counterpart in the reference. */
int phx_lw_autotune(phx_lw* h, int nk1, const int* k1_list, double* out4);
/* Mesh-line batching of the outer points: the points of a phx_lw_run call
 * that share a mesh line (equal j1, j2 of k1 = sum_d j_d/m_d b_d) are
 * processed together, up to 64 per batch, when the line holds at least
 * min_points of them; cacheElPh then streams gW once per line and factors
 * e^{ik1.Re} = e^{2 pi i j0 n0/m0} e^{2 pi i (j1 n1/m1 + j2 n2/m2)} over the
 * lattice coordinates n_d of R_e.  The remaining points keep the k1-batched
 * GEMM.  0 turns it off (the default).  Results agree with the unbatched run
 * to rounding.  No counterpart in the reference, whose cacheElPh
 * (interaction_elph.cpp:229-448) sums over R_e once per outer point. */
int phx_lw_set_line_batching(phx_lw* h, int min_points);
/* Panels ((k1, b1) pairs) per launch of the separable transform; 0 (the
 * default) = as many as keep one chunk's g3 rows of the split last stage
 * within 24 GiB, between 16 and 128.  The stage buffers hold one chunk:
 * (lines * mz + groups * my * mz) * nm * nw complex values per panel for T1 /
 * T2, and mesh points * max(nm nw, 2 nb2max nm) complex FP32 for g3.  No
 * counterpart in the reference. */
int phx_lw_set_panel_chunk(phx_lw* h, int panels);
/* Last stage of the separable transform as two kernels (on, the default) or
 * one fused kernel (off).  Split: an x-stage kernel writes the kept rows' g3
 * to device memory, (panel chunk) x (y, z lines) x m0 x nm nw complex FP32,
 * and an epilogue kernel reads them; both use far fewer registers and LDS
 * than the fused kernel, so more workgroups hide the memory latency.  Applies
 * to PHX_FP32 with nm, nw <= 16, nm nw <= 192, an x mesh <= 64 and at most 24
 * x-groups; elsewhere the fused kernel runs.  The g3 values are the same
 * either way and the linewidths agree to FP64 rounding.
 * on = 1 (default) runs the lean epilogue kernel where it applies (fixed
 * Gaussian smearing, at most 8 bands per inner point; it also takes the
 * learned screen's slot masks), on = 2 the first epilogue kernel always;
 * both give the same linewidths.  on = 3 moves the electron rotation
 * g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2) (FP64, the epilogue's own
 * arithmetic on the same FP32 g3) into the x-stage kernel, which writes g'
 * (nb2 x nm complex FP64 per kept row, slot rows of nb2max x nm) instead of
 * g3, and the lean epilogue starts from g'; where the lean epilogue does not
 * apply it runs as on = 1.  The linewidths agree with on = 1 to FP64
 * rounding.  No counterpart in the reference. */
int phx_lw_set_split_last_stage(phx_lw* h, int on);
/* Split last stage: run the epilogue kernel of each panel chunk on a second
 * stream of the handle, concurrently with the next chunk's transform stages
 * (on = 1), or after them on the handle's stream (on = 0, the default).  With
 * on = 1 the chunks alternate between two sets of g3 rows and row masks
 * (twice the device memory of those); the learned screen's passes always run
 * in turn.  The linewidths are the same either way.  Takes effect for batches
 * prepared after the call.  No counterpart in the reference. */
int phx_lw_set_epilogue_overlap(phx_lw* h, int on);
/* Split last stage, PHX_FP32 without the learned screen: run the y-stage of
 * the separable transform inside the x-stage kernel (on = 1, the default),
 * so that the x-stage slabs T2 stay in LDS instead of going through device
 * memory, or as its own kernel (on = 0).  Not with split mode 3.  The g3
 * values, and so the linewidths, are the same either way.  No counterpart
 * in the reference. */
int phx_lw_set_fused_y_stage(phx_lw* h, int on);
/* With timing on: *on = 1 when every timed split launch ran the y-stage inside
 * the x-stage kernel (the x-stage times and bytes then cover both stages).
 * No counterpart in the reference (a measurement call). */
int phx_lw_last_stage_fused_y(phx_lw* h, int* on);
/* With timing on, the launches of phx_lw_stage_times that ran split:
 * out4 = {x-stage kernel ms, epilogue kernel ms, launches, sum over the rows
 * the epilogue computed of their band counts}.  No counterpart in the reference. */
int phx_lw_last_stage_split_times(phx_lw* h, double* out4);
/* With timing on and the last stage split: out2 = {bytes the x-stage kernel
 * moved, bytes of its output the epilogue read}.  The x-stage kernel reads
 * the x-stage slabs (slots computed x x-groups x nm nw complex FP32) and
 * writes g3 (kept rows x nm nw complex FP32); with the rotation in the
 * x-stage (split mode 3) it writes g' instead (kept bands x nm complex FP64)
 * and reads U2 (kept bands x nw complex FP64); with the y-stage inside it
 * reads the y-stage input T1 (panels x lines x mz x nm nw complex FP32)
 * instead of the slabs.  No counterpart in the
 * reference (a measurement call). */
int phx_lw_last_stage_split_bytes(phx_lw* h, double* out2);
/* ---- Learned q-block screen (separable transform with the energy screen on).
 * No counterpart in the reference, which evaluates every (k1, k2) pair
 * (el_scattering.cpp:161-386).  A q-block is one panel (k1, b1) times one
 * (jy, jz) line of inner points.  A 16 -> 64 -> 64 -> 1 MLP with BF16 operands
 * and FP32 accumulation predicts ln of the block's contribution to the
 * linewidths of b1's degenerate group, as a residual on the block's Gaussian
 * baseline ln(sum over its rows of e^{-x^2}), x a row's closest approach to
 * resonance in units of sigma.  Blocks with a row within 3 sigma of
 * resonance are always computed; the others are skipped when their predicted
 * contribution, raised by the 99.9th-percentile under-prediction seen in
 * training, stays below threshold times the panel's predicted total.
 * audit_frac of the skipped blocks are recomputed exactly, and a block found
 * above 10 x the threshold makes the batch recompute every skipped block.
 *
 * Collect training samples: runs the exact last stage on a random fraction
 * of the blocks inside the energy screen (resonant and tail blocks alike)
 * for the given outer points (all when k1_list is null) and appends their
 * features and log contributions.  The linewidths are not touched.  Returns
 * the total sample count.  The training's under-prediction margin is taken
 * over the samples without a row within 3 sigma (the blocks the screen may
 * skip). */
int phx_lw_screen_collect(phx_lw* h, int nk1, const int* k1_list, double sample_frac,
                          unsigned long long seed);
/* Train the network on the collected samples on the GPU (Adam, MSE on the
 * standardized log contribution); loss_first_last receives the full-set loss
 * before and after training. */
int phx_lw_screen_train(phx_lw* h, int epochs, int batch, double lr, unsigned long long seed,
                        double* loss_first_last);
int phx_lw_set_learned_screen(phx_lw* h, int on, double threshold, double audit_frac,
                              unsigned long long seed);
/* out8: blocks inside the energy screen, learned skips, audited, batches that
 * fell back, max audited contribution / threshold, mean audited contribution /
 * threshold, training samples, under-prediction margin (ln units). */
int phx_lw_screen_stats(phx_lw* h, double* out8);
int phx_lw_screen_reset_stats(phx_lw* h);
/* Read (set = 0) or install (set != 0) the phx_mlp_num_params() FP32 network parameters. */
int phx_lw_screen_params(phx_lw* h, float* params, int set);
int phx_mlp_num_params(void);
/* Loss, gradients and outputs of the network for n samples (X: n x 16,
 * row-major; t: n targets), computed by the GPU forward / backward kernels. */
int phx_mlp_gradients(const float* params, const float* X, const float* t, int n, float* grads,
                      float* loss, float* y);
/* Timing of the separable transform's last stage (the fused kernel, or the
 * x-stage and epilogue kernels when it runs split), with HIP events
 * around each of its launches while timing is on: out4 = {milliseconds,
 * launches, rows computed (after the screens), FLOPs of its mesh-transform
 * GEMMs for those rows = 8 * rows * groups * nm * nw}. */
int phx_lw_stage_times(phx_lw* h, double* out4);
/* A plain device-to-device copy of `bytes` (float4 grid-stride kernel), timed
 * over `reps` launches: the HBM reference of the roofline. */
int phx_stream_copy(unsigned long long bytes, int reps, double* ms_per_copy);
/* Per-kernel timing with HIP events recorded on the handle's stream around
 * every batch; turning it on (or off) clears the accumulators. */
int phx_lw_set_timing(phx_lw* h, int on);
/* out5 = {cacheElPh ms, coupling ms, collect ms, batches timed,
 *         algorithmic FLOPs of the timed coupling launches}. */
int phx_lw_kernel_times(phx_lw* h, double* out5);
int phx_lw_sync(phx_lw* h);
/* HIP stream the handle enqueues on (hipStream_t). */
void* phx_lw_stream(phx_lw* h);
/* Copy the handle's gW to host memory (nRe*nRp*nm*nw*nw complex). */
int phx_lw_copy_gw(phx_lw* h, double* out);
int phx_lw_destroy(phx_lw* h);

/* ---------------------------------------------------------------- phonon linewidths
 * PhElScatteringMatrix::builder (phel_scattering.cpp:43-363): the phonon
 * linewidths from el-ph scattering, the second caller of the coupling.
 *   prob: recip, mesh (electron k mesh), the electron Hamiltonian, the force
 *         constants (+ long-range term), nCalc / temperatures / chemPots,
 *         sigma, winMin / winMax (electron window), smearing; its coupling
 *         fields are unused;
 *   coupling: the el-ph interaction (phx_elph_create*, phx_elph_create_from_file),
 *         which must outlive the handle;
 *   q_mesh: the Gamma-centred phonon mesh; every q point and mode is a state.
 * For every active k1 and every q3: k2 = k1 + q3 (all bands), |g|^2
 * symmetrized, Gamma(q3, nu) += |g|^2 f1 (1 - f1) delta(e1 - e2 + w) / (T N_k) pi. */
typedef struct phx_phel phx_phel;
int phx_phel_setup(phx_phel** h, const phx_problem* prob, phx_elph* coupling, const int* q_mesh);
int phx_phel_num_active(phx_phel* h);
/* outer points k1_list (active indices; nk1 < 0: all); returns the (k1, q3) pairs done */
long long phx_phel_run(phx_phel* h, int nk1, const int* k1_list);
/* out: nCalc x (nq x nm), (q, nu) row-major, Ry; average: degeneracy averaging
 * over the phonon modes of each q, row 0 only (scattering.cpp:1168-1207) */
int phx_phel_get_linewidths(phx_phel* h, double* out, int average);
/* phonon frequencies nq x nm of the q mesh */
int phx_phel_get_phonons(phx_phel* h, double* omega);
int phx_phel_reset(phx_phel* h);
int phx_phel_destroy(phx_phel* h);
/* Kernel timing of phx_phel_run with HIP events on the coupling's stream
 * (one host synchronization per batch of 32 outer points while on); switching
 * it on or off clears the totals.  No counterpart in the reference. */
int phx_phel_set_timing(phx_phel* h, int on);
/* out6 = {cacheElPh ms, coupling ms, symmetrize + rates ms, coupling launches,
 * cacheElPh launches, R_ph-transform FLOPs of the coupling launches}. */
int phx_phel_kernel_times(phx_phel* h, double* out6);

/* Host twin of the synthetic gW generator (bit-identical to the device). */
int phx_synth_gw_host(double* gW, int nw, int nm, int nRp, int nRe, const double* decay,
                      unsigned long long seed);

#ifdef __cplusplus
}
#endif

#endif /* PHX_ELPH_H */
