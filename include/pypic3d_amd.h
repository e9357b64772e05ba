/*
 * pypic3d_amd.h -- C ABI of the MI355X (gfx950) PIC hot-path library
 * libpypic3d_hip.so.
 *
 * Every entry point takes plain device pointers, POD descriptors and a
 * hipStream_t passed as void*.  No torch types cross this boundary.  Return
 * codes: 0 success, <0 invalid argument (the Python layer raises ValueError),
 * >0 a hipError_t (RuntimeError).  All work is stream-ordered; no entry point
 * synchronises the host except pic_poisson_cg (which must read its residual)
 * and none allocates device memory: scratch comes in through `workspace`
 * arguments sized by the matching *_workspace_bytes query.
 *
 * Layout (one tile covering the domain, i.e. the reference's single-device
 * layout with the leading (1,1,1) tile axes dropped):
 *   field component   float32[(Nx+2g)][(Ny+2g)][(Nz+2g)], z fastest
 *   particle planes   float32[S][P] per component (x, y, z, vx, vy, vz),
 *                     uint8 active[S][P]; plane index s*P + p
 * Reference citations are PyPIC3D/<file>:<line>.
 */
#ifndef PYPIC3D_AMD_H
#define PYPIC3D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_ABI_VERSION 12
#define PIC_MAX_SPECIES 16

/* Scalar run description: StaticParameters + DynamicParameters fields the hot
 * path reads (parameters.py:14-53) plus the tile-local grid lines
 * (grids.py:80-111) reduced to origin/spacing, computed on the host in FP64
 * exactly as the reference builds them. */
typedef struct PicDesc {
    int32_t n[3];          /* physical cells Nx, Ny, Nz                       */
    int32_t g;             /* guard cells                                     */
    int32_t dims[3];       /* n + 2g                                          */
    int32_t shape_factor;  /* 1 = CIC, 2 = TSC                                */
    int32_t relativistic;  /* relativistic Boris (boris.py:60-124)            */
    int32_t bc[3];         /* field BC: 0 periodic, 1 conducting              */
    int32_t pbc[3];        /* particle BC: 0 periodic, 1 reflect, 2 absorb    */
    int32_t variant;       /* kernel variant selector (tuning); 0 = default   */
    double d[3];           /* dx, dy, dz                                      */
    double wind[3];        /* x_wind, y_wind, z_wind                          */
    double org_c[3];       /* tiled "center" line: first coordinate           */
    double sp_c[3];        /*                      line[1] - line[0]          */
    double org_v[3];       /* tiled "vertex" line                             */
    double sp_v[3];
    double dt, C, eps, mu, alpha;
    int32_t pusher;        /* 0 Boris (relativistic per the flag above),      */
    int32_t pad1;          /* 1 Higuera-Cary (higuera_cary.py:57-113)         */
} PicDesc;

/* SpeciesConfig (particles/particle_class.py:5-14).  count may be 0 and
 * slots may be 0 (an empty set, as in the reference's particle-free runs):
 * its planes may then be null and the particle entry points do nothing. */
typedef struct PicSpecies {
    int32_t count;
    int32_t slots;                    /* P: slots per species               */
    double charge[PIC_MAX_SPECIES];
    double mass[PIC_MAX_SPECIES];
    double weight[PIC_MAX_SPECIES];
    uint8_t update_x[PIC_MAX_SPECIES][3];
    uint8_t update_u[PIC_MAX_SPECIES][3];
} PicSpecies;

/* SoA particle planes (TiledParticles x, u, active; particle_class.py:17-31). */
typedef struct PicParticles {
    float* x[3];
    float* u[3];
    uint8_t* active;
} PicParticles;

int pic_abi_version(void);

/* ---- field utilities -------------------------------------------------- */
/* Fill n floats with value (graph-capture safe replacement for memset). */
int pic_fill_f32(float* ptr, int64_t n, float value, void* stream);
/* update_tiled_ghost_cells for one tile (ghost_cells.py:142-215, 615-650).
 * bc_type 0 uses desc->bc, 1 uses desc->pbc (ghost_cells.py:33-39). */
int pic_halo_refresh(const PicDesc* desc, float* field, int bc_type, void* stream);
/* fold_tiled_ghost_cells (ghost_cells.py:218-316, 703-736). */
int pic_halo_fold(const PicDesc* desc, float* field, int bc_type, void* stream);
/* apply_tiled_zero_boundary / apply_tiled_constant_boundary on one axis
 * (ghost_cells.py:344-386, 653-700); both refresh afterwards. */
int pic_zero_boundary(const PicDesc* desc, float* field, int axis, void* stream);
int pic_constant_boundary(const PicDesc* desc, float* field, int axis, void* stream);
/* digital_filter / bilinear_filter (filters.py:234-291): out interior <- 3x3x3
 * VALID correlation of in; out ghosts <- in ghosts.  in != out. */
int pic_filter_digital(const PicDesc* desc, const float* in, float* out, double alpha, void* stream);
int pic_filter_bilinear(const PicDesc* desc, const float* in, float* out, void* stream);

/* ---- Yee solver (solvers/first_order_yee.py) --------------------------- */
/* B_out interior = B_in - (dt/2) curl E with E neighbours read as refreshed
 * ghosts (first_order_yee.py:96-142); then B_out ghosts refreshed.  The
 * do_filter branch (:145-159) is pic_filter_digital + pic_halo_refresh. */
int pic_yee_update_b(const PicDesc* desc, const float* const E[3], const float* const B_in[3],
                     float* const B_out[3], void* stream);
/* E_out interior = E_in + dt (C^2 curl B - J/eps) (first_order_yee.py:12-74);
 * then E_out ghosts refreshed.  Filter and conducting walls are separate calls. */
int pic_yee_update_e(const PicDesc* desc, const float* const E_in[3], const float* const B[3],
                     const float* const J[3], float* const E_out[3], void* stream);

/* PML (boundary_conditions/PML.py): ADE memory of the six curl derivatives
 * (interior (Nx, Ny, Nz) float32 arrays, updated in place; order dEz/dy,
 * dEy/dz, dEx/dz, dEz/dx, dEy/dx, dEx/dy for B and the same with B for E)
 * and the damping profiles sigma_x, sigma_y, sigma_z in the field layout. */
typedef struct PicPml {
    float* mem[6];
    const float* sigma[3];
    /* optional (NULL: unused): sigma_line[a][i] = sigma[a] at index i of axis
     * a (dims[a] entries) for a profile that varies along its own axis only
     * over the interior, as PML.py's are; the kernels then read the line
     * instead of the field-shaped profile (results are identical). */
    const float* sigma_line[3];
} PicPml;
/* pic_yee_update_b / _e with apply_tiled_pml_to_b_curl / _e_curl
 * (PML.py:270-359; first_order_yee.py:54-59, 133-138) applied to the
 * derivatives: b = exp(-sigma dt) with the full dt, m' = b m + (b-1) d,
 * d' = d + m'. */
int pic_yee_update_b_pml(const PicDesc* desc, const float* const E[3], const float* const B_in[3],
                         float* const B_out[3], const PicPml* pml, void* stream);
int pic_yee_update_e_pml(const PicDesc* desc, const float* const E_in[3], const float* const B[3],
                         const float* const J[3], float* const E_out[3], const PicPml* pml, void* stream);
/* stretch_spatial_derivative on six interior derivative arrays in place, in
 * the memory order above (the derivative part of apply_tiled_pml_to_*_curl). */
int pic_pml_stretch(const PicDesc* desc, float* const deriv[6], const PicPml* pml, void* stream);

/* ---- particles --------------------------------------------------------- */
/* particle_push (pusher/particle_push.py:13-175): gather E+E_ext, B+B_ext on
 * the Yee-staggered grids, Boris / relativistic Boris / Higuera-Cary per
 * desc->pusher and desc->relativistic (particle_push.py:109-132); writes u_out (may
 * alias u).  E_ext/B_ext may be NULL (zero external field). */
int pic_push(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
             const float* const E[3], const float* const B[3],
             const float* const E_ext[3], const float* const B_ext[3],
             float* const u_out[3], void* stream);
/* update_tiled_particle_positions (particle_tile_communication.py:82-99) with
 * step dt*dt_scale; x_out may alias in->x. */
int pic_advance_positions(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                          double dt_scale, float* const x_out[3], void* stream);
/* refresh_tiled_particle_tiles on one device (particle_tile_communication.py:
 * 41-59, 297-375): wrap/reflect/absorb per axis, zero x,u of inactive slots.
 * In place on io. */
int pic_particle_boundaries(const PicDesc* desc, const PicSpecies* sp, PicParticles* io, void* stream);
/* Raw (unfolded) deposits into zero-initialised ghosted arrays:
 * Esirkepov_current deposit (deposition/Esirkepov.py:105-331),
 * J_from_rhov deposit (deposition/J_from_rhov.py:88-200),
 * compute_rho deposit (deposition/rho.py:72-153). */
int pic_deposit_esirkepov(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                          float* const J[3], void* stream);
int pic_deposit_direct(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                       float* const J[3], void* stream);
int pic_deposit_rho(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                    float* rho, void* stream);
/* Tile-ordered versions of the two deposits above for particles sorted by
 * pic_tile_keys (ranges from pic_tile_ranges, tile 4 or 8): one block per
 * (species, tile) accumulates in an LDS box around the tile and adds it into
 * the arrays once.  Same results up to FP32 summation order. */
int pic_deposit_direct_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* const J[3],
                             int32_t tile, const int32_t* starts, void* stream);
int pic_deposit_rho_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* rho, int32_t tile,
                          const int32_t* starts, void* stream);
/* Fused particle phase of time_loop_electrodynamic's Esirkepov branch
 * (evolve.py:36-43, 70-79): gather + push, raw Esirkepov deposit into J (zeroed
 * by the caller), x += v dt, particle boundaries.  Writes out (must not alias
 * in).  desc->variant picks one of PIC_FUSED_VARIANTS launch shapes (slots per
 * thread x resident-block target) for the 3-D first-order case. */
#define PIC_FUSED_VARIANTS 6
int pic_fused_push_esirkepov(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                             const float* const E[3], const float* const B[3],
                             const float* const E_ext[3], const float* const B_ext[3],
                             PicParticles* out, float* const J[3], void* stream);

/* Tile-ordered variant for particles sorted by pic_tile_keys (3-D first-order
 * case only): one block per (species, tile) slot range [starts[b], starts[b+1])
 * from pic_tile_ranges accumulates its J contributions in an LDS box around the
 * tile and adds the box into J once.  Same results as
 * pic_fused_push_esirkepov up to FP32 summation order.  E+E_ext and B+B_ext
 * around the tile are staged in LDS for the gather.  tile: 4 or 8 cells per
 * axis; rounds: 2 or 4 particle slots per thread per pass; min_blocks: 1, 3
 * or 4 resident blocks per CU requested from the register allocator.
 * On an exact grid (every node origin + a*spacing an FP32 number, e.g. integer
 * or dyadic spacings and origins) a thread takes `rounds` consecutive sorted
 * slots and sums equal-cell deposits in registers; particles its fast loop does
 * not cover are listed in workspace (1 + count*slots int32: workspace[0] is
 * the list length) and finished by a second kernel in the same call.
 * desc->variant = 1 selects the strided kernel on every grid (comparison).
 * pic_tile_exact_grid returns 1 when desc is such an exact grid. */
int pic_tile_exact_grid(const PicDesc* desc);
int pic_fused_push_esirkepov_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                   const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                   const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                   const int32_t* starts, int32_t rounds, int32_t min_blocks, int32_t* workspace,
                                   void* stream);
/* pic_fused_push_esirkepov_tiled on particles re-binned without their
 * gather (pic_tile_rebin with out == NULL): slot i of the sorted order reads
 * in's slot gather[i] (perm of that re-bin; starts index the sorted order)
 * and writes out's slot i, so out holds the sorted particles after the push.
 * Results equal pic_permute_particles(in, gather) followed by
 * pic_fused_push_esirkepov_tiled: the particles bit for bit, J up to FP32
 * summation order.  Exact grids only (pic_tile_exact_grid, desc->variant !=
 * 1); -1 otherwise.  gather may be NULL for an empty set. */
int pic_fused_push_esirkepov_tiled_gather(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                          const int32_t* gather, const float* const E[3], const float* const B[3],
                                          const float* const E_ext[3], const float* const B_ext[3],
                                          PicParticles* out, float* const J[3], int32_t tile, const int32_t* starts,
                                          int32_t rounds, int32_t min_blocks, int32_t* workspace, void* stream);
/* Fused particle phase of time_loop_electrodynamic's direct branch
 * (evolve.py:44-67) for tile-sorted particles on an exact 3-D first-order grid
 * whose cell size equals the node spacing (desc->d == desc->sp_c): gather +
 * push, x += v dt/2, particle boundaries, J_from_rhov deposit into J (zeroed
 * by the caller; fold, refresh and filter stay with the caller), x += v dt/2,
 * particle boundaries.  Writes out (must not alias in); same results as
 * pic_push, pic_advance_positions, pic_particle_boundaries,
 * pic_deposit_direct, pic_advance_positions, pic_particle_boundaries up to
 * FP32 rounding (an ulp in a velocity where the compiler fuses the push's
 * multiply-adds differently) and summation order in J.  tile 4 or 8;
 * min_blocks 1 (each thread's four slots loaded and stored with 16-byte
 * accesses) or 4 (tile 4: two slots per thread with 8-byte accesses, which
 * fits 128 VGPRs for 4 resident blocks per CU, about 15 % faster at 256^3 x
 * 32 ppc; tile 8: slots stored one by one); workspace as for
 * pic_fused_push_esirkepov_tiled. */
int pic_fused_push_direct_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                const int32_t* starts, int32_t min_blocks, int32_t* workspace, void* stream);
/* pic_fused_push_direct_tiled reading slot i of the sorted order from in's
 * slot gather[i] (the perm of a pic_tile_rebin called with out == NULL), as
 * pic_fused_push_esirkepov_tiled_gather does for the Esirkepov branch:
 * results equal pic_permute_particles(in, gather) followed by
 * pic_fused_push_direct_tiled, the particles bit for bit, J up to FP32
 * summation order. */
int pic_fused_push_direct_tiled_gather(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                       const int32_t* gather, const float* const E[3], const float* const B[3],
                                       const float* const E_ext[3], const float* const B_ext[3], PicParticles* out,
                                       float* const J[3], int32_t tile, const int32_t* starts, int32_t min_blocks,
                                       int32_t* workspace, void* stream);

/* ---- cell sort (particle_initialization.py:89-102, 140-160) ------------- */
/* key = s*(N+1) + (ix*Ny+iy)*Nz+iz with ix = clip(floor((x+W/2)/dx),0,Nx-1);
 * inactive slots get key s*(N+1) + N. */
int pic_cell_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                  int32_t* keys, void* stream);
/* Tile-major key for the tile-ordered fused kernel (a reordering for
 * locality only; results do not depend on particle order): the cell
 * (ix, iy, iz) = clip(floor((x + W/2) * (1/d)), 0, N-1) in FP64 (the product
 * with the reciprocal can differ from pic_cell_keys' quotient for a position
 * within an ulp of a cell face), tiles of tile^3 cells in x-major order,
 * cells x-major inside a tile, inactive slots last in their species.
 * count * slots < 2^31. */
int64_t pic_tile_count(const PicDesc* desc, int32_t tile);
int pic_tile_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t tile, int32_t* keys,
                  void* stream);
/* starts[s*(tiles+1) + t] = first sorted slot of tile t of species s (t =
 * tiles: the inactive tail); starts[count*(tiles+1)] = count*slots. */
int pic_tile_ranges(const PicDesc* desc, const PicSpecies* sp, int32_t tile, const int32_t* keys_sorted,
                    int32_t* starts, void* stream);
/* Counting re-bin of the particles into tile-key order (the order
 * pic_tile_keys + a sort produce, with the order inside one key unspecified):
 * out receives the particles of in grouped by ascending tile key, starts the
 * slot ranges of pic_tile_ranges, perm the source slot of every output slot.
 * A counting pass (keys aggregated per 1024-slot chunk in an LDS hash table,
 * one atomic per distinct key and chunk), a scan of the counts, a pass that
 * writes perm from each chunk's reserved positions, and a gather of the
 * particles through perm.  Replaces pic_tile_keys + pic_sort_permutation +
 * pic_permute_particles + pic_tile_ranges.  lookahead: the keys are taken at
 * the predicted position x + lookahead * u (FP32; axes the species does not
 * move stay put), 0 for the current position: a caller that re-sorts every k
 * steps passes about k dt / 2, so the particles are closest to their sorted
 * cells in the middle of the interval instead of at its start (any order
 * gives the same results up to FP32 summation order).  out == NULL skips the
 * gather: the caller applies perm itself (pic_permute_particles, or the next
 * step's pic_fused_push_esirkepov_tiled_gather / pic_fused_push_direct_tiled_gather).  workspace: the query below. */
int64_t pic_tile_rebin_workspace_bytes(const PicDesc* desc, const PicSpecies* sp, int32_t tile);
int pic_tile_rebin(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t tile,
                   double lookahead, PicParticles* out, int32_t* starts, int32_t* perm, void* workspace,
                   int64_t workspace_bytes, void* stream);
int64_t pic_sort_workspace_bytes(int64_t n);
/* Stable radix sort of (key, slot) pairs: perm = stable argsort(keys);
 * keys_sorted receives the sorted keys. */
int pic_sort_permutation(const int32_t* keys, int32_t* keys_sorted, int32_t* perm, int64_t n,
                         int32_t key_bits, void* workspace, int64_t workspace_bytes, void* stream);
/* out planes[i] = in planes[perm[i]] (7 planes: x,y,z,ux,uy,uz,active); n < 2^31. */
int pic_permute_particles(const PicParticles* in, PicParticles* out, const int32_t* perm,
                          int64_t n, void* stream);

/* ---- electrostatics (solvers/electrostatic_yee.py) ---------------------- */
int64_t pic_cg_workspace_bytes(const PicDesc* desc);
/* solve_poisson_with_conjugate_gradient (electrostatic_yee.py:71-156) as
 * mixed-precision iterative refinement: phi and the residual
 * r = rho/eps + lapl(phi) in FP64, each correction solved by CG with FP32
 * vectors and FP64 dot products.  Stops when the FP64 ||r|| <= max(tol,
 * rtol*||r0||), when a refinement pass no longer halves ||r||, or after
 * max_iter CG iterations in total.  phi in/out (its boundaries applied as
 * _apply_tiled_phi_constant_boundaries does).  Synchronises the host once
 * per refinement pass and once per 32 iterations to read the convergence
 * flag.  iterations (total CG iterations) and residual_norm (final FP64
 * ||r||) optional. */
int pic_poisson_cg(const PicDesc* desc, const float* rho, float* phi, double tol, double rtol,
                   int32_t max_iter, void* workspace, int64_t workspace_bytes,
                   int32_t* iterations, double* residual_norm, void* stream);
/* _centered_tiled_electrostatic_gradient (electrostatic_yee.py:212-246):
 * E interior = -centred grad phi, ghosts refreshed (phi must be refreshed). */
int pic_efield_from_phi(const PicDesc* desc, const float* phi, float* const E[3], void* stream);

/* ---- diagnostics (utils.py:108-202) ------------------------------------ */
int64_t pic_energy_workspace_bytes(const PicDesc* desc, const PicSpecies* sp);
/* compute_energy (utils.py:108-187): out[3] (device, FP64) = electric,
 * magnetic, kinetic energy, accumulated in FP64. */
int pic_energy(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
               const float* const E[3], const float* const B[3], double* out,
               void* workspace, int64_t workspace_bytes, void* stream);
/* pic_energy without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Its partial sums live
 * in a static 24 KiB device array per GPU (held until the process ends); it
 * calls no allocator.  Every call on one GPU shares that array: calls that
 * overlap on one GPU race.  Returns -2 (with a message on stderr) if a call
 * would need more partial-sum space than the array holds.  The stream must
 * belong to the current device. */
int pic_energy_noalloc(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                       const float* const E[3], const float* const B[3], double* out, void* stream);
/* compute_total_momentum (utils.py:190-202): out[0] (device, FP64) =
 * sum over active slots of |v| mass weight.  workspace: at least 8192 bytes. */
int pic_total_momentum(const PicSpecies* sp, const PicParticles* in, double* out, void* workspace,
                       int64_t workspace_bytes, void* stream);
/* Total charge: out[0] (device, FP64) = sum over the interior of rho dx dy dz
 * (the charge diagnostic; the reference sums rho with jnp.sum where its tests
 * check charge, e.g. single_particle_pipeline_test.py:429-431).  workspace:
 * at least 8192 bytes. */
int pic_total_charge(const PicDesc* desc, const float* rho, double* out, void* workspace, int64_t workspace_bytes,
                     void* stream);

/* ---- FP64 path ------------------------------------------------------------
 * The reference runs in FP64 (__main__.py:186-187).  These entry points take
 * double field components (same layout) and double particle planes and
 * follow the reference's expressions in FP64, one thread per slot, any
 * dimensionality and shape order.  Same semantics as the FP32 entry point of
 * the same name without the suffix; J/rho are raw deposits into
 * zero-initialised arrays (pic_fill_f64). */
typedef struct PicParticlesF64 {
    double* x[3];
    double* u[3];
    uint8_t* active;
} PicParticlesF64;
int pic_halo_refresh_f64(const PicDesc* desc, double* field, int bc_type, void* stream);
int pic_halo_fold_f64(const PicDesc* desc, double* field, int bc_type, void* stream);
int pic_zero_boundary_f64(const PicDesc* desc, double* field, int axis, void* stream);
int pic_constant_boundary_f64(const PicDesc* desc, double* field, int axis, void* stream);
int pic_filter_digital_f64(const PicDesc* desc, const double* in, double* out, double alpha, void* stream);
int pic_filter_bilinear_f64(const PicDesc* desc, const double* in, double* out, void* stream);
int pic_yee_update_b_f64(const PicDesc* desc, const double* const E[3], const double* const B_in[3],
                         double* const B_out[3], void* stream);
int pic_yee_update_e_f64(const PicDesc* desc, const double* const E_in[3], const double* const B[3],
                         const double* const J[3], double* const E_out[3], void* stream);
int pic_push_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, const double* const E[3],
                 const double* const B[3], const double* const E_ext[3], const double* const B_ext[3],
                 double* const u_out[3], void* stream);
int pic_advance_positions_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double dt_scale,
                              double* const x_out[3], void* stream);
int pic_particle_boundaries_f64(const PicDesc* desc, const PicSpecies* sp, PicParticlesF64* io, void* stream);
int pic_deposit_esirkepov_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in,
                              double* const J[3], void* stream);
int pic_deposit_direct_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double* const J[3],
                           void* stream);
int pic_deposit_rho_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double* rho,
                        void* stream);
/* wrap_periodic_position (boundary_conditions/grid_and_stencil.py:14-35) on n
 * positions in place: mod(x + W/2, W) - W/2, a position at +W/2 kept there
 * (evaluated in FP64; the FP32 variant rounds the result to FP32). */
int pic_wrap_periodic(float* x, int64_t n, double wind, void* stream);
int pic_wrap_periodic_f64(double* x, int64_t n, double wind, void* stream);
/* workspace: at least 24576 bytes */
int pic_energy_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, const double* const E[3],
                   const double* const B[3], double* out, void* workspace, int64_t workspace_bytes, void* stream);

/* ---- batched ensembles ----------------------------------------------------
 * Independent simulations of uneven grid size and particle count advance
 * together, each with time_loop_electrodynamic's Esirkepov-branch semantics
 * (evolve.py:16-103).  Particle planes and field components of all members
 * are slices of shared batch arrays in the single-simulation layout.  The
 * reference runs one simulation per process; this has no direct counterpart
 * there. */
typedef struct PicBatchSim {
    PicDesc desc;
    PicSpecies species;
    int64_t particle_offset;   /* first slot of this member in the batch planes   */
    int64_t field_offset;      /* first element of this member's field arrays     */
} PicBatchSim;

/* Host-side validation of a table (periodic fields and particles, alpha = 1,
 * slices inside the batch arrays). */
int pic_batch_check(const PicBatchSim* sims, int32_t n_sims, int64_t particle_slots, int64_t field_elems);
/* Fused particle phase of every member: `sims` and `work` are device arrays;
 * work holds n_work (member, block) pairs covering each member's slots in
 * blocks of 256 * rounds (rounds 2, 4 or 8); `grid` persistent blocks walk
 * the list.  J (batch arrays) must be zero on entry.  out must not alias in. */
int pic_batch_fused_push_esirkepov(const PicBatchSim* sims, const int32_t* work, int32_t n_work, int32_t rounds,
                                   int32_t grid, const PicParticles* in, const float* const E[3],
                                   const float* const B[3], PicParticles* out, float* const J[3], void* stream);
/* Grid phase of every member, one block per member: J fold + refresh
 * (particle codes), B_half = B - dt/2 curl E_in, E_out = E_in + dt (C^2 curl
 * B_half - J/eps), B = B_half - dt/2 curl E_out, each refreshed. */
int pic_batch_field_step(const PicBatchSim* sims, int32_t n_sims, const float* const E_in[3], float* const B[3],
                         float* const B_half[3], float* const J[3], float* const E_out[3], void* stream);

/* ---- adjoint of the electrodynamic Esirkepov step (periodic fields and
 * particles; the reference has no gradient code -- these are the reverse-mode
 * transposes of evolve.py:16-103 as computed by the entry points above).
 * Adjoint arrays are FP64.  "interior" arrays are (Nx, Ny, Nz), z fastest;
 * "ghosted" arrays have the field layout. ------------------------------- */
int pic_fill_f64(double* ptr, int64_t n, double value, void* stream);
/* ghosted[r] = interior[periodic partner of r]: transpose of fold + refresh
 * (ghost_cells.py:142-289) for the adjoint of J. */
int pic_periodic_expand_f64(const PicDesc* desc, const double* interior, double* ghosted, void* stream);
/* interior[partner(r)] += ghosted[r]: transpose of the ghost refresh read by
 * the gather, for the adjoint of E and B. */
int pic_periodic_fold_add_f64(const PicDesc* desc, const double* ghosted, double* interior, void* stream);
/* Transpose of update_B(E) -> update_E(B_h, J) -> update_B(E', filter)
 * (first_order_yee.py:12-162; evolve.py:86-97) on interior arrays:
 * in: lamE, lamB = adjoints of the step's E', B'; out (in place): their
 * adjoints with respect to the step's E, B through the grid update, and
 * lamJ = adjoint of the folded J.  work: six interior scratch arrays. */
int pic_adjoint_yee(const PicDesc* desc, double* const lamE[3], double* const lamB[3], double* const lamJ[3],
                    double* const work[6], void* stream);
/* Transpose of the particle phase (particle_push.py gather + pusher,
 * Esirkepov.py deposit, particle_tile_communication.py:82-99 position update)
 * at the step's start state `state` and fields E, B (+ external):
 * lam_x, lam_v in = adjoints of x', v'; out (in place) = adjoints of x, v.
 * lamJ: interior adjoint of J (read at each deposit node's periodic partner,
 * the transpose of fold + refresh); lamE, lamB: interior FP64 arrays to which
 * the adjoint of the gathered fields is added (atomics at the gather nodes'
 * periodic partners, the transpose of the ghost refresh).  min_blocks: kept
 * for compatibility, not used (the first-order path's two kernels fit two
 * waves per SIMD unforced).  workspace: 1 + count*slots int32 (first-order
 * particles that move two or more cells are finished by a third, general
 * pass).  scratch: 7*count*slots doubles between the first-order path's two
 * kernels (push-Jacobian products and a done flag per slot); may be null for
 * second-order shapes. */
int pic_adjoint_particles(const PicDesc* desc, const PicSpecies* sp, const PicParticles* state,
                          const float* const E[3], const float* const B[3], const float* const E_ext[3],
                          const float* const B_ext[3], const double* const lamJ[3], double* const lam_x[3],
                          double* const lam_v[3], double* const lamE[3], double* const lamB[3], int32_t min_blocks,
                          int32_t* workspace, double* scratch, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYPIC3D_AMD_H */
