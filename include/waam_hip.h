/*
 * waam_hip.h — C ABI of libwaam_hip.so, the gfx950 implementation of waam_twin's coupled
 * physics step.
 *
 * The reference (THEDARKDEACON/waam_solver, Python + Taichi) has no FFI: its hot path is a set
 * of @ti.kernel functions called positionally from Python with Taichi fields. Each entry point
 * below replaces one of those kernels (reference file:line given per entry) and keeps its
 * positional parameter order, with four mechanical changes:
 *   - a field argument becomes a device pointer (float* or uint8_t* for flags);
 *   - FLAG_* arguments are dropped: flag values are fixed by this ABI (WAAM_FLAG_*, the
 *     reference's grid.py:63-66 values);
 *   - the grid sizes nx, ny, nz are followed by the field layout (WAAM_LAYOUT_*);
 *   - 0-d scratch fields (norm_buf, vol_acc, ...) become device pointers, and every entry
 *     takes a trailing hipStream_t (passed as void*, NULL = the null stream).
 *
 * Conventions
 *   - Layout: every field of a grid uses one layout. WAAM_LAYOUT_XYZ is the reference's
 *     C-order (nx, ny, nz), k fastest; WAAM_LAYOUT_ZYX is C-order (nz, ny, nx), i fastest.
 *     Populations are 19 such arrays back to back. Flags are uint8 with FLUID=0, SOLID=1,
 *     GAS=2, IFACE=4.
 *   - All launches are asynchronous on the given stream. Entries that return a host value
 *     (*_sync) synchronise that stream.
 *   - Return 0 on success or a negative WAAM_ERR_* code; waam_last_error() gives a message
 *     (thread-local). No exception crosses the ABI.
 *   - The library never allocates or frees caller memory. Entries that reduce over the grid
 *     take a caller-provided device workspace of at least waam_scratch_bytes() bytes.
 *   - Not re-entrant on the same field set; distinct grids on distinct streams are safe.
 */
#ifndef WAAM_HIP_H
#define WAAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WAAM_ABI_VERSION 5

enum { WAAM_FLAG_FLUID = 0, WAAM_FLAG_SOLID = 1, WAAM_FLAG_GAS = 2, WAAM_FLAG_IFACE = 4 };
/* the moving lid of the lid-driven cavity (waam_cavity_step); not a flag of the coupled step */
enum { WAAM_CAVITY_LID = 3 };
enum { WAAM_LAYOUT_XYZ = 0, WAAM_LAYOUT_ZYX = 1 };
/* Population storage of the fused step: fp32 populations, (f - w_q) * 2^15 in binary16
 * (saturating at +-65504), or f - w_q in bfloat16 (round to nearest even on store, fp32
 * arithmetic in between). */
enum { WAAM_POP_F32 = 0, WAAM_POP_F16 = 1, WAAM_POP_BF16 = 2 };
enum { WAAM_OK = 0, WAAM_ERR_INVALID = -1, WAAM_ERR_HIP = -2 };

int waam_abi_version(void);
const char *waam_last_error(void);
/* Bytes of device workspace that reducing entries need (arc weight total, telemetry,
 * deposition budget, surface height). */
int64_t waam_scratch_bytes(void);

/* ---- initialisation: init_grid kernels.py:62-120 -------------------------------------- */
int waam_init_grid(float *f, float *rho, float *ux, float *uy, float *uz, float *T, float *H,
                   float *f_l, float *phi, uint8_t *flags, float T_ambient, float rho0,
                   float cp_rho, int nz_solid, int plate_i0, int plate_i1, int plate_j0,
                   int plate_j1, int nx, int ny, int nz, int layout, void *stream);

/* ---- arc heat: inject_arc_heat kernels.py:443-510, inject_goldak_heat :548-622 --------- */
/* norm_out (device float, may be NULL) receives the weight total; scratch is the workspace. */
int waam_inject_arc_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                         float arc_i, float arc_j, float arc_k, float Q_w, float sigma, float dt,
                         float dx3, float eta, float penetration_cells, int enable_surface_weight,
                         int nx, int ny, int nz, int layout, void *scratch, void *stream);
int waam_inject_goldak_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                            float arc_i, float arc_j, float arc_k, float Q_w, float dt,
                            float dx3, float eta, float travel_sign, float ff, float fr,
                            float a_front, float a_rear, float b_axis, float c_axis,
                            float penetration_cells, int enable_surface_weight, int nx, int ny,
                            int nz, int layout, void *scratch, void *stream);
/* surface_height_at kernels.py:625-641 (max qualifying k; result as float in *out). */
int waam_surface_height_at(const float *phi, const uint8_t *flags, float *out, int i0, int j0,
                           int nz_fallback, int nx, int ny, int nz, int layout, void *stream);

/* ---- thermal ----------------------------------------------------------------------------- */
/* clamp_enthalpy_ceiling_scalar kernels.py:1953-1984 */
int waam_clamp_enthalpy_ceiling_scalar(float *H, const uint8_t *flags, float cp_rho,
                                       float T_solidus, float T_liquidus, float T_vapor_cap,
                                       float L_rho, int64_t n, void *stream);
/* clamp_enthalpy_floor_scalar kernels.py:1937-1950 */
int waam_clamp_enthalpy_floor_scalar(float *H, float cp_rho, const uint8_t *flags,
                                     float T_floor, int64_t n, void *stream);
/* advect_diffuse_temperature kernels.py:1611-1682, Jacobi form: reads H_in, writes H_out
 * (H_in == H_out is rejected). */
int waam_advect_diffuse_temperature(const float *H_in, float *H_out, const float *T,
                                    const float *ux, const float *uy, const float *uz,
                                    const uint8_t *flags, float alpha_lu, float dt,
                                    float cp_rho, int nx, int ny, int nz, int layout, void *stream);
/* apply_thermal_boundary_losses kernels.py:2030-2082 */
int waam_apply_thermal_boundary_losses(float *H, const float *T, const uint8_t *flags,
                                       float T_amb, float h_conv, float eps_rad,
                                       int enable_conv, int enable_rad, float cp_rho, float dt,
                                       float dx, float sigma_sb, int nx, int ny, int nz, int layout,
                                       void *stream);
/* update_phase kernels.py:1519-1556 */
int waam_update_phase(const float *H, float *T, float *f_l, float cp_rho, float L_rho,
                      float T_solidus, float T_liquidus, int64_t n, void *stream);
/* update_T_max kernels.py:3016-3032 */
int waam_update_T_max(const float *T, float *T_max, const uint8_t *flags, int64_t n,
                      void *stream);
/* update_cooling_rate kernels.py:2013-2027 */
int waam_update_cooling_rate(const float *T, float *T_prev, float *dT_dt, const uint8_t *flags,
                             float dt_phys, int64_t n, void *stream);
/* update_time_above_T kernels.py:3261-3283 */
int waam_update_time_above_T(const float *T, const uint8_t *flags, float *t800, float *t1100,
                             float *tsol, float dt, float T_800, float T_1100, float T_solidus,
                             int64_t n, void *stream);

/* ---- VOF ----------------------------------------------------------------------------------- */
/* advect_phi kernels.py:2320-2361 (+ _vof_face_flux :2275-2317) */
int waam_advect_phi(float *phi_dst, const float *phi_src, const float *ux, const float *uy,
                    const float *uz, const uint8_t *flags, int nx, int ny, int nz, int layout, void *stream);
/* reinitialize_phi kernels.py:2364-2391 */
int waam_reinitialize_phi(float *phi, const uint8_t *flags, int64_t n, void *stream);
/* apply_contact_angle_phi_bc kernels.py:933-991, Jacobi form (phi_in != phi_out); cos/sin of
 * the contact angle are passed precomputed. */
int waam_apply_contact_angle_phi_bc(const float *phi_in, float *phi_out, const uint8_t *flags,
                                    float cos_t, float sin_t, int nx, int ny, int nz, int layout,
                                    void *stream);
/* update_flags_from_phi kernels.py:2394-2419 */
int waam_update_flags_from_phi(const float *phi, const float *f_l, uint8_t *flags, int nz_solid,
                               int nx, int ny, int nz, int layout, void *stream);
/* sync_phi_liquid_fraction kernels.py:3484-3504 */
int waam_sync_phi_liquid_fraction(const float *phi, float *f_l, const uint8_t *flags, int64_t n,
                                  void *stream);
/* solidify_cooled_metal kernels.py:2422-2450 */
int waam_solidify_cooled_metal(const float *T, float *f_l, float *phi, uint8_t *flags, float *ux,
                               float *uy, float *uz, float T_solidus, int zero_velocity,
                               int64_t n, void *stream);
/* remelt_hot_solid_scalar kernels.py:2489-2520 */
int waam_remelt_hot_solid_scalar(const float *T, float *H, float *f_l, float *phi,
                                 uint8_t *flags, float cp_rho, float T_solidus,
                                 float T_liquidus, float L_rho, int64_t n, void *stream);
/* solidify_trailing_pool_scalar kernels.py:2573-2617 */
int waam_solidify_trailing_pool_scalar(float *T, float *H, float *f_l, float *phi,
                                       uint8_t *flags, float *ux, float *uy, float *uz,
                                       float cp_rho, float arc_i, float arc_j, float arc_k,
                                       float dir_x, float dir_y, float dir_z,
                                       float lookback_cells, float T_freeze, float T_solidus,
                                       int nx, int ny, int nz, int layout, void *stream);

/* ---- forces ---------------------------------------------------------------------------- */
/* compute_csf_tension kernels.py:858-926 */
int waam_compute_csf_tension(const float *phi, const uint8_t *flags, float *Fx, float *Fy,
                             float *Fz, float gamma_lu, int enable_wetting, float cos_t,
                             float sin_t, int nx, int ny, int nz, int layout, void *stream);
/* compute_curvature_field kernels.py:3353-3376 */
int waam_compute_curvature_field(const float *phi, const uint8_t *flags, float *kappa_out,
                                 int nx, int ny, int nz, int layout, void *stream);
/* compute_marangoni_force kernels.py:2158-2217 */
int waam_compute_marangoni_force(const float *T, const float *phi, const float *f_l, float *Fx,
                                 float *Fy, float *Fz, const uint8_t *flags, float dgamma_dT,
                                 int nx, int ny, int nz, int layout, void *stream);
/* add_buoyancy kernels.py:2729-2753 */
int waam_add_buoyancy(const float *T, float *Fz, const float *f_l, const uint8_t *flags,
                      float g_lu, float beta, float T_ref, float rho_ref, int64_t n,
                      void *stream);
/* add_hydrostatic_gravity kernels.py:2756-2772 */
int waam_add_hydrostatic_gravity(float *Fz, const float *f_l, const uint8_t *flags,
                                 float rho_ref, float g_lu, int64_t n, void *stream);
/* apply_arc_pressure kernels.py:3038-3083 */
int waam_apply_arc_pressure(float *Fz, const uint8_t *flags, const float *phi, float arc_i,
                            float arc_j, float arc_k, float sigma, float pressure_pa, float dt,
                            float dx, float rho_ref, int nx, int ny, int nz, int layout, void *stream);
/* clamp_body_force_magnitude kernels.py:3286-3309 */
int waam_clamp_body_force_magnitude(float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                    float F_max_lu, int64_t n, void *stream);
/* apply_droplet_impact_pressure kernels.py:1142-1173 */
int waam_apply_droplet_impact_pressure(float *Fz, const uint8_t *flags, const float *phi,
                                       const float *f_l, float arc_i, float arc_j, float arc_k,
                                       float drop_radius, float impact_pa, float dt, float dx,
                                       float rho_ref, int nx, int ny, int nz, int layout, void *stream);

/* ---- LBM ----------------------------------------------------------------------------------- */
/* collide_mrt cumulant_kernel.py:28-209 */
int waam_collide_mrt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const float *Fx, const float *Fy, const float *Fz,
                     const float *f_l, const uint8_t *flags, float omega_s, float omega_b,
                     float C_darcy, int nx, int ny, int nz, int layout, void *stream);
/* collide_srt kernels.py:2778-2885 */
int waam_collide_srt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const float *Fx, const float *Fy, const float *Fz,
                     const float *f_l, const uint8_t *flags, float tau, float omega,
                     float dt_lu, float C_darcy, int nx, int ny, int nz, int layout, void *stream);
/* collide_srt_variable_tau kernels.py:2888-2965 */
int waam_collide_srt_variable_tau(const float *f_src, float *f_dst, float *rho, float *ux,
                                  float *uy, float *uz, const float *Fx, const float *Fy,
                                  const float *Fz, const float *f_l, const uint8_t *flags,
                                  const float *tau_field, float C_darcy, int nx, int ny, int nz, int layout,
                                  void *stream);
/* stream kernels.py:2971-3010 (pull, x periodic, y/z clamped, bounce-back) */
int waam_stream(const float *f_src, float *f_dst, const uint8_t *flags, int nx, int ny, int nz, int layout,
                void *stream);
/* clamp_velocity_mach kernels.py:3312-3335 */
int waam_clamp_velocity_mach(float *ux, float *uy, float *uz, const uint8_t *flags,
                             float u_max_lu, int64_t n, void *stream);

/* ---- deposition, tracers ------------------------------------------------------------------- */
/* feed_wire_surface kernels.py:278-400 — one call, serial-order (i, j, k sweep) semantics;
 * vol_acc/real_acc are device floats (the reference's 0-d fields). */
int waam_feed_wire_surface(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                           float *T, float *rho, float arc_i, float arc_j, float arc_k,
                           float footprint_r, float droplet_radius, float target_vol,
                           float T_drop, float cp_rho, float L_rho, float rho0, float *vol_acc,
                           float *real_acc, float cell_vol, int nx, int ny, int nz, int layout,
                           void *scratch, void *stream);
/* The droplet placement loop of solvers/coupled_step.py:126-151 (up to 10 feed_wire_surface
 * calls, footprint growing by 1.5 cells to at most 16) in one launch, with no host round trip.
 * footprint_r and drop_vol are the host's double values; result (device, 2 floats) receives the
 * placed volume and the number of calls made. */
int waam_deposit_droplet(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                         float *T, float *rho, float arc_i, float arc_j, float arc_k,
                         double footprint_r, float droplet_radius, double drop_vol,
                         float T_drop, float cp_rho, float L_rho, float rho0, float *vol_acc,
                         float *real_acc, float cell_vol, float *result, int nx, int ny, int nz, int layout,
                         void *scratch, void *stream);
/* feed_wire_momentum_impact kernels.py:1176-1233 */
int waam_feed_wire_momentum_impact(float *f_src, const float *rho, float *ux, float *uy,
                                   float *uz, const uint8_t *flags, const float *f_l,
                                   float arc_i, float arc_j, float arc_k, float droplet_radius,
                                   float vx_lu, float vy_lu, float vz_lu, int nx, int ny, int nz, int layout,
                                   void *stream);
/* inject_tracers kernels.py:3182-3204 (counter-hash RNG in place of ti.random) */
int waam_inject_tracers(float *pos, int32_t *active, int32_t *head, int max_tracers,
                        float arc_x, float arc_y, float arc_z, float sigma_m, int spawn_count,
                        uint32_t seed, void *stream);
/* advect_tracers kernels.py:3206-3255 */
int waam_advect_tracers(float *pos, int32_t *active, const float *ux, const float *uy,
                        const float *uz, const float *f_l, const uint8_t *flags, float dx,
                        float dt, int max_tracers, int nx, int ny, int nz, int layout, void *stream);

/* ---- temperature-dependent material path ---------------------------------------------------
 * Piecewise-linear property tables (reference gpu_tables.py:11-66: at most 8 knots per
 * property; an empty table becomes one knot at 300 K holding the constant) and the per-cell
 * property fields they fill. Tables and reference values are passed by value. */
#define WAAM_MAX_KNOTS 8
typedef struct waam_material_tables {
    float cp_T[WAAM_MAX_KNOTS], cp_V[WAAM_MAX_KNOTS];
    float k_T[WAAM_MAX_KNOTS], k_V[WAAM_MAX_KNOTS];
    float mu_T[WAAM_MAX_KNOTS], mu_V[WAAM_MAX_KNOTS];
    float dgamma_T[WAAM_MAX_KNOTS], dgamma_V[WAAM_MAX_KNOTS];
    int32_t n_cp, n_k, n_mu, n_dgamma;
    float cp_fallback, k_fallback, mu_fallback, dgamma_fallback;
} waam_material_tables;
/* scalars of refresh_thermal_properties (kernels.py:1710-1745): rho, dt, dx, the Marangoni
 * prefactor force_scale (scale_F/dx), the constant-property values for SOLID/GAS cells and the
 * Marangoni calibration scale */
typedef struct waam_property_refs {
    float rho, dt, dx, force_scale;
    float cp_rho_ref, alpha_lu_ref, dgamma_lu_ref, tau_ref, marangoni_scale;
} waam_property_refs;

/* refresh_thermal_properties kernels.py:1710-1772 (n cells; use_tables 0 gives the constants) */
int waam_refresh_thermal_properties(const float *T, float *cp_rho_field, float *alpha_lu_field,
                                    float *dgamma_lu_field, float *tau_field,
                                    const waam_material_tables *tables,
                                    const waam_property_refs *refs, int use_tables,
                                    const uint8_t *flags, int64_t n, void *stream);
/* advect_diffuse_temperature_variable kernels.py:1775-1836 (Jacobi, H_in -> H_out) */
int waam_advect_diffuse_temperature_variable(const float *H_in, float *H_out, const float *T,
                                             const float *ux, const float *uy, const float *uz,
                                             const uint8_t *flags, const float *alpha_lu_field,
                                             const float *cp_rho_field, float dt, int nx, int ny,
                                             int nz, int layout, void *stream);
/* update_phase_variable_cp kernels.py:1839-1863 */
int waam_update_phase_variable_cp(const float *H, float *T, float *f_l, const float *cp_rho_field,
                                  float L_rho, float T_solidus, float T_liquidus, float H_sol,
                                  float H_liq, int64_t n, void *stream);
/* clamp_enthalpy_floor kernels.py:1937-1950 */
int waam_clamp_enthalpy_floor(float *H, const float *cp_rho_field, const uint8_t *flags,
                              float T_floor, int64_t n, void *stream);
/* clamp_enthalpy_ceiling_variable_cp kernels.py:1987-2010 */
int waam_clamp_enthalpy_ceiling_variable_cp(float *H, const uint8_t *flags,
                                            const float *cp_rho_field, float H_liq,
                                            float T_liquidus, float T_vapor_cap, int64_t n,
                                            void *stream);
/* compute_marangoni_force_variable kernels.py:2220-2269 */
int waam_compute_marangoni_force_variable(const float *T, const float *phi, const float *f_l,
                                          float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                          const float *dgamma_lu_field, int nx, int ny, int nz,
                                          int layout, void *stream);
/* remelt_hot_solid kernels.py:2454-2486 */
int waam_remelt_hot_solid(const float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                          const float *cp_rho_field, float T_solidus, float T_liquidus,
                          float L_rho, int64_t n, void *stream);
/* solidify_trailing_pool kernels.py:2524-2570 */
int waam_solidify_trailing_pool(float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                                float *ux, float *uy, float *uz, const float *cp_rho_field,
                                float arc_i, float arc_j, float arc_k, float dir_x, float dir_y,
                                float dir_z, float lookback_cells, float T_freeze,
                                float T_solidus, int nx, int ny, int nz, int layout, void *stream);

/* ---- full-tier surface forces (arc_k is accepted and unused, as in the reference) ---------- */
/* apply_vapor_recoil kernels.py:995-1030: Fz += -(p/(rho dx)) dt^2/dx min(T/T_ref, 3)
 * exp(-r_ij^2 / (2 sigma^2)) on non-SOLID/GAS cells with |phi(k+1) - phi(k-1)| >= 0.05 */
int waam_apply_vapor_recoil(float *Fz, const float *T, const float *phi, const uint8_t *flags,
                            float arc_i, float arc_j, float arc_k, float sigma, float recoil_pa,
                            float T_ref, float dt, float dx, float rho_ref, int nx, int ny,
                            int nz, int layout, void *stream);
/* apply_vapor_recoil_clausius_clapeyron kernels.py:1052-1090: p = C_acc P_ref
 * exp(min(L/R (1/T_boil - 1/T), 12)) above T_boil */
int waam_apply_vapor_recoil_clausius_clapeyron(float *Fz, const float *T, const float *phi,
                                               const uint8_t *flags, float arc_i, float arc_j,
                                               float arc_k, float sigma, float P_ref_Pa,
                                               float T_boil_K, float L_vapor_J_kg,
                                               float R_spec_J_kgK, float C_acc, float dt, float dx,
                                               float rho_ref, int nx, int ny, int nz, int layout,
                                               void *stream);
/* apply_gas_shear_stress kernels.py:1093-1139: tangential (Fx, Fy) shear away from the arc
 * axis where |grad phi| >= 0.08 */
int waam_apply_gas_shear_stress(float *Fx, float *Fy, const float *phi, const uint8_t *flags,
                                float arc_i, float arc_j, float arc_k, float sigma,
                                float tau_peak_pa, float dt, float dx, float rho_ref, int nx,
                                int ny, int nz, int layout, void *stream);

/* ---- Lorentz force of the welding current: kernels.py:1236-1513, driven by
 * physics/weld_forces.py:196-281 (electric potential by weighted Jacobi, plate grounded) --- */
/* elec_build_sigma kernels.py:1236-1251: 1e-8 in GAS, sigma_liquid where f_l > 0.55 */
int waam_elec_build_sigma(float *sigma, const float *f_l, const uint8_t *flags,
                          float sigma_liquid, float sigma_solid, int64_t n, void *stream);
/* elec_clear_source kernels.py:1254-1257 */
int waam_elec_clear_source(float *source, int64_t n, void *stream);
/* elec_inject_arc_source kernels.py:1260-1286 (cells of the slab range only) */
int waam_elec_inject_arc_source(float *source, float arc_i, float arc_j, float arc_k,
                                float sigma_cells, float current_A, float dx, int nx, int ny,
                                int nz, int layout, void *stream);
/* elec_normalize_source kernels.py:1289-1301: source *= I / max(sum source dx^3, 1e-12) */
int waam_elec_normalize_source(float *source, float target_current_A, float dx, int64_t n,
                               void *scratch, void *stream);
/* elec_init_ground kernels.py:1304-1312: phi = 0 */
int waam_elec_init_ground(float *phi, int64_t n, void *stream);
/* elec_jacobi_step kernels.py:1315-1356 */
int waam_elec_jacobi_step(const float *phi_in, float *phi_out, const float *sigma,
                          const float *source, const uint8_t *flags, float dx, float omega_j,
                          int nz_solid, int nx, int ny, int nz, int layout, void *stream);
/* elec_l1_diff kernels.py:1384-1396: out (device, 2 floats) = (sum |a - b|, sum |b|) */
int waam_elec_l1_diff(const float *a, const float *b, float *out, int64_t n, void *scratch,
                      void *stream);
/* elec_compute_J kernels.py:1359-1381 */
int waam_elec_compute_J(float *Jx, float *Jy, float *Jz, const float *phi, const float *sigma,
                        const uint8_t *flags, float dx, int nx, int ny, int nz, int layout,
                        void *stream);
/* elec_bin_axial_current kernels.py:1399-1418: bins is nz x n_bins fp32 (slab-major) */
int waam_elec_bin_axial_current(const float *Jz, float *bins, float arc_i, float arc_j, float dx,
                                float bin_dr_cells, int n_bins, int nx, int ny, int nz,
                                int layout, void *stream);
/* elec_prefix_bins kernels.py:1421-1432 */
int waam_elec_prefix_bins(float *bins, int n_bins, int nz, void *stream);
/* elec_B_axisymmetric kernels.py:1435-1477 */
int waam_elec_B_axisymmetric(float *Bx, float *By, float *Bz, const float *bins, float arc_i,
                             float arc_j, float dx, float bin_dr_cells, float mu0, int n_bins,
                             int nx, int ny, int nz, int layout, void *stream);
/* apply_lorentz_JxB kernels.py:1480-1513 */
int waam_apply_lorentz_JxB(float *Fx, float *Fy, float *Fz, const float *Jx, const float *Jy,
                           const float *Jz, const float *Bx, const float *By, const float *Bz,
                           const float *f_l, const uint8_t *flags, float dt, float dx,
                           float rho_ref, int64_t n, void *stream);

/* ---- moving window: shift_simulation_window_x kernels.py:2620-2723 ------------------------ */
/* Every listed field (both population buffers included) becomes new[i] = old[i + n_shift] for
 * i < nx - n_shift (the reference's ascending serial loop), and the last n_shift columns are
 * reset: T = T_amb, H = cp_rho T_amb, T_max = T_prev = T_amb, f_l = u = F = dT_dt = times = 0,
 * rho = rho0, populations w_q rho0, SOLID with phi 1 below nz_solid and GAS with phi 0 above.
 * The per-cell material fields are shifted and not reset, as in the reference. */
int waam_shift_simulation_window_x(int n_shift, float *f_a, float *f_b, float *T, float *H,
                                   float *f_l, float *phi, uint8_t *flags, float *T_max,
                                   float *T_prev, float *dT_dt, float *time_above_800,
                                   float *time_above_1100, float *time_above_solidus, float *rho,
                                   float *ux, float *uy, float *uz, float *Fx, float *Fy,
                                   float *Fz, float *cp_rho_field, float *alpha_lu_field,
                                   float *dgamma_lu_field, float *tau_field, float T_amb,
                                   float cp_rho, float rho0, int nz_solid, int nx, int ny, int nz,
                                   int layout, void *stream);

/* ---- bead height for the CTWD update: measure_bead_height_m physics/electrical_stickout.py:39-50
 * out (device int32): the largest k of a SOLID cell with k >= nz_solid, or -1 when there is none
 * (the reference computes it on the host from every flag). */
int waam_bead_top_k(const uint8_t *flags, int nz_solid, int32_t *out, int nx, int ny, int nz,
                    int layout, void *stream);

/* ---- lid-driven cavity --------------------------------------------------------------------
 * No counterpart in the reference: its stream (kernels.py:2982-3010) has bounce-back only and
 * its buffer choreography moves no momentum between fluid cells. One step of the plain D3Q19
 * cavity: every FLUID cell pulls from its neighbours (all axes periodic; walls are cells),
 * half-way bounce-back off other cells, plus 6 w_q (e_q . u_w) off WAAM_CAVITY_LID cells
 * (Ladd's moving wall, wall density 1), then collide_mrt's map (cumulant_kernel.py:28-209)
 * without force or Darcy damping. Writes f_dst, rho and u of FLUID cells only. */
int waam_cavity_step(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const uint8_t *flags, float omega_s, float omega_b, float uwx,
                     float uwy, float uwz, int nx, int ny, int nz, int layout, void *stream);

/* ---- the differentiation form: adjoint of the thermal sub-step ----------------------------
 * No counterpart in the reference (it has no derivatives). The forward map is the coupled
 * step's thermal chain (solvers/coupled_step.py:97-160) with frozen flags and zero velocity:
 * Gaussian2D arc injection, enthalpy ceiling, conduction from the step's starting temperature,
 * boundary losses, enthalpy floor, phase map, ceiling, phase map. Parameters: the arc
 * efficiency eta, a factor loss_scale on both loss coefficients (h_conv0, eps_rad0 are the
 * unscaled ones) and a factor sigma / sigma0 on the arc radius (cells). */
typedef struct waam_adj_consts {
    float cp_rho, L_rho, T_s, T_l;       /* phase map */
    float T_cap, T_floor, alpha_lu;      /* ceiling temperature, floor temperature, conduction */
    float T_amb, h_conv0, eps_rad0, sigma_sb, dt, dx, loss_scale;
    float arc_i, arc_j, arc_k, sigma, sigma0, Q_w, eta, dx3, pen;
    int32_t enable_cap, enable_loss, enable_arc, enable_conv, enable_rad, surface_weight;
} waam_adj_consts;
/* out2 (device, 2 doubles) = sum over weighted non-GAS cells of wg wdep and of wg wdep r^2 */
int waam_adj_arc_sums(const uint8_t *flags, const float *f_l, const waam_adj_consts *c,
                      double *out2, int nx, int ny, int nz, int layout, void *scratch,
                      void *stream);
/* lam = grad_T dT/dH + grad_fl df_l/dH at enthalpy H (0 on GAS cells) */
int waam_adj_terminal(const float *H, const float *grad_T, const float *grad_fl,
                      const uint8_t *flags, const waam_adj_consts *c, float *lam, int64_t n,
                      void *stream);
/* One step back: lam_out = dL/dH_n from lam_next = dL/dH_{n+1}, given the step's enthalpy at
 * its start (H_n), after the arc (A), before the floor (E) and after it (F), the starting
 * f_l and the step's arc sums; adds the step's dL/d(eta, loss_scale, sigma scale) to grads
 * (device, 3 doubles). lamC is an n-cell work buffer. */
int waam_adj_thermal_step(const float *lam_next, float *lam_out, const float *H_n, const float *A,
                          const float *E, const float *F, const float *f_l_n,
                          const uint8_t *flags, const double *arc_sums,
                          const waam_adj_consts *c, double *grads, float *lamC, int nx, int ny,
                          int nz, int layout, void *scratch, void *stream);
/* Adjoint of compute_marangoni_force (kernels.py:2158-2217; waam_compute_marangoni_force) with
 * phi, f_l and the flags held, for the force adjoint lam = dL/dF (three n-cell fp32 fields):
 * T_adj (n doubles) += dL/dT, and s_cell[c] (n doubles) = lam_c . F_c / dgamma, whose sum is
 * dL/d(dgamma). G is a work buffer of 3n doubles. */
int waam_adj_marangoni(const float *T, const float *phi, const float *f_l, const uint8_t *flags,
                       const float *lam_x, const float *lam_y, const float *lam_z, float dgamma,
                       double *T_adj, double *s_cell, double *G, int nx, int ny, int nz,
                       int layout, void *stream);

/* Adjoint of collide_srt (kernels.py:2815-2885; waam_collide_srt) with f_l and the flags held:
 * for lam = dL/d(post-collision populations) (19 q-major n-cell fp32 arrays), lam_f (19n doubles)
 * = dL/d(f_src) and lam_F (3n doubles) = dL/d(Fx, Fy, Fz); 0 on SOLID and GAS cells, which the
 * collision does not write. */
int waam_adj_collide_srt(const float *f_src, const float *Fx, const float *Fy, const float *Fz,
                         const float *f_l, const uint8_t *flags, const float *lam, float omega,
                         float C_darcy, double *lam_f, double *lam_F, int64_t n, void *stream);

/* Adjoint of stream (kernels.py:2991-3010; waam_stream), flags held: lam_in (19 q-major n-cell
 * doubles) = dL/d(f_src) for lam_out = dL/d(f_dst); 0 for inputs no output copies. */
int waam_adj_stream(const double *lam_out, double *lam_in, const uint8_t *flags, int nx, int ny,
                    int nz, int layout, void *stream);

/* ---- adjoint of the coupled step with frozen flags and phi (binary64 adjoint state) --------
 * The step's thermal chain with enthalpy advection (kernels.py:1611-1682 and the chain of
 * waam_adj_thermal_step), the body forces the temperature and liquid fraction drive
 * (Marangoni kernels.py:2158-2217, hydrostatic and buoyancy kernels.py:2232-2268, the force
 * clamp), the MRT collision (cumulant_kernel.py:59-209) and the Mach clamp, as the coupled step
 * (solvers/coupled_step.py:97-409) runs them with the flags and phi held; with frozen flags the
 * stream never feeds a collided cell. */
/* collision: lam_f (19 q-major), lam_u (3 arrays, after the Mach clamp at u_cap) -> lam_fin
 * (19), lam_F (3, the force the collision saw), lam_fl += (the drag's f_l); FLUID/IFACE cells */
int waam_cadj_collide_mrt(const float *f_src, const float *Fx, const float *Fy, const float *Fz,
                          const float *f_l, const uint8_t *flags, const double *lam_f,
                          const double *lam_u, double omega_s, double omega_b, double C_darcy,
                          double u_cap, double *lam_fin, double *lam_F, double *lam_fl, int64_t n,
                          void *stream);
/* force clamp at the raw force Fr (F_cap), hydrostatic and buoyancy terms: lam_F -> lam_Fraw
 * (3 arrays); lam_T += and lam_fl += */
int waam_cadj_forces(const float *Frx, const float *Fry, const float *Frz, const float *T,
                     const float *f_l, const uint8_t *flags, const double *lam_F, double F_cap,
                     double g_lu, double beta, double T_l, int enable_hydro, double *lam_Fraw,
                     double *lam_T, double *lam_fl, int64_t n, void *stream);
/* Marangoni force with binary64 lam (3 arrays): T_adj += dL/dT; G is 3n doubles of workspace */
int waam_cadj_marangoni(const float *phi, const float *f_l, const uint8_t *flags,
                        const double *lam, double dgamma, double *G, double *T_adj, int nx, int ny,
                        int nz, int layout, void *stream);
/* phase map at H: lam_H += lam_T dT/dH + lam_fl df_l/dH (0 on GAS) */
int waam_cadj_phase(const float *H, const uint8_t *flags, const waam_adj_consts *ac,
                    const double *lam_T, const double *lam_fl, double *lam_H, int64_t n,
                    void *stream);
/* thermal chain with advection: lam_next = dL/dH_{n+1} -> lam_out = dL/dH_n, lam_u (3 arrays) =
 * dL/du_n; H_n, A (after the arc), B (after the first ceiling), E (after the losses), F (after
 * the floor), f_l_n and u_n of the forward; grads += dL/d(eta, s_h, s_sigma); lamC and lamB are
 * n doubles of workspace each */
int waam_cadj_thermal_step(const double *lam_next, double *lam_out, double *lam_u,
                           const float *H_n, const float *A, const float *B, const float *E,
                           const float *F, const float *f_l_n, const float *ux, const float *uy,
                           const float *uz, const uint8_t *flags, const double *arc_sums,
                           const waam_adj_consts *ac, double *grads, double *lamC, double *lamB,
                           int nx, int ny, int nz, int layout, void *scratch, void *stream);

/* ---- export-side fields and reductions -------------------------------------------------- */
/* compute_vorticity_magnitude kernels.py:3380-3419: |curl u| / dt on non-GAS cells, 0 on GAS */
int waam_compute_vorticity_magnitude(const float *ux, const float *uy, const float *uz,
                                     const uint8_t *flags, float *vort_out, float dx, float dt,
                                     int nx, int ny, int nz, int layout, void *stream);
/* out (device float) = max |(fx, fy, fz)| over n cells (force_linf_lu,
 * physics/force_diagnostics.py:19-24) */
int waam_max_norm3(const float *fx, const float *fy, const float *fz, float *out, int64_t n,
                   void *stream);
/* ProbeRecorder.record_step export/probes.py:95-113: out[2p], out[2p+1] = T, f_l at linear cell
 * cells[p] (in the fields' layout); a negative entry (probe outside the window) gives NaN. The
 * caller points out at a row of a device ring of samples and reads the ring back in one copy. */
int waam_probe_gather(const float *T, const float *f_l, const int64_t *cells, int n_probes,
                      int64_t n_cells, float *out, void *stream);

/* ---- telemetry: telemetry_pool_reduce kernels.py:3424-3468 ------------------------------- */
/* out (device, 8 doubles): n_liquid, T_peak, cool, u2_max, i_sum, i_min, i_max, T_global. */
int waam_telemetry_pool_reduce(const float *T, const float *f_l, const float *phi,
                               const float *dT_dt, const float *ux, const float *uy,
                               const float *uz, double *out, int nx, int ny, int nz, int layout,
                               void *scratch, void *stream);

/* ---- the fused step ------------------------------------------------------------------------
 * One coupled timestep (reference solvers/coupled_step.py:73-464, scalar-material path) in seven
 * launches, with results identical to the one-kernel-per-reference-kernel sequence above.
 * Fields that a kernel both reads around a cell and writes come in pairs (*_in read, *_out
 * written) that the caller swaps after the step. The pull stream of step n is deferred into the
 * collide kernel of step n+1 ("pending stream"): it is materialised only where the collide of
 * step n+1 does not overwrite it, using the flags of step n (sflags_prev). See
 * waam_stream_pending for materialising it on demand. */
typedef struct waam_step_fields {
    /* populations in the form pop_dtype names (float, or 16-bit words of f - w_q) */
    void *f_src, *f_dst;
    float *rho, *ux, *uy, *uz;
    float *H_in, *H_out, *T_in, *T_out, *fl_in, *fl_out;
    float *phi, *phi_tmp;
    uint8_t *flags_in, *flags_out, *sflags_prev, *sflags_next;
    float *Fx, *Fy, *Fz;
    float *T_max, *T_prev, *dT_dt, *t800, *t1100, *tsol;
    float *tracer_pos;
    int32_t *tracer_active;
    /* two int32 work lists of nx*ny*nz entries each (compacted FLUID/IFACE cell indices) */
    int32_t *list_thermal, *list_collide;
    /* waam_quiet_bytes(nx, ny, nz, layout) bytes: per row segment (the 64 cells of one
     * wavefront along the contiguous axis) whether it ended the last two steps quiet (all GAS,
     * phi 0) or cold (SOLID below the solidus and the HAZ thresholds), the list of the segments
     * a steady step sweeps and which of them the flag-update and stream sweeps visit, and the
     * pull's per-lane cell masks; written by every step, contents need no initial value (the
     * size is queried, not a layout) */
    uint8_t *quiet;
    /* temperature-dependent material (params.use_tables; NULL otherwise): cp_rho_prev holds
     * the previous step's refresh_thermal_properties cp*rho (kernels.py:1710-1772), read by the
     * ceiling that follows arc injection; the step writes its own refresh into cp_rho,
     * alpha_lu, dgamma_lu and tau for every cell (cp_rho_prev and cp_rho then swap roles) */
    float *cp_rho_prev, *cp_rho, *alpha_lu, *dgamma_lu, *tau;
    void *scratch;
    int32_t max_tracers, nx, ny, nz;
    int32_t layout; /* WAAM_LAYOUT_XYZ or WAAM_LAYOUT_ZYX, shared by every field */
    int32_t pop_dtype; /* WAAM_POP_* of f_src and f_dst */
} waam_step_fields;

typedef struct waam_step_params {
    /* switches */
    int32_t is_welding, collision /* 0 MRT, 1 SRT, 2 SRT, tau per cell */, arc_kind /* 0 Gaussian, 1 Goldak */;
    int32_t arc_applied /* arc + ceiling already applied this step (droplet steps) */;
    int32_t pending_stream, forces_from_fields, store_forces;
    int32_t enable_vof, enable_wetting, enable_csf, enable_heat_loss, enable_conv, enable_rad;
    int32_t enable_cap, enable_hydro, enable_freeze, enable_trailing, zero_velocity;
    int32_t surface_weight, nz_solid, col_i, col_j, use_arc_k_override, use_k_robot;
    /* 1: clear the forces of every cell the collision skips; 0: only of those the previous
     * fused step collided (sflags_prev FLUID/IFACE), valid when every other cell's forces are
     * already zero (the step before was a fused step that did not take forces from the fields) */
    int32_t force_reset_all;
    /* 0: the thermal sweep builds the FLUID/IFACE list and a second launch consumes it.
     * 1 or 2: the list already holds this step's FLUID/IFACE cells, appended by the previous
     * waam_step_fused call's last kernel (valid when nothing changed flags_in since that call);
     * 1 runs the surface cells in the sweep's launch, 2 in a launch of their own (the shared
     * launch lowers the sweep's occupancy: it pays on short sweeps, not on long ones). */
    int32_t thermal_list_ready;
    /* full-tier surface forces (physics/weld_forces.py:120-161): vapour recoil (clausius: the
     * Clausius-Clapeyron form, else the constant-pressure form) and shielding-gas shear */
    int32_t enable_recoil, recoil_clausius, enable_gas_shear;
    /* 1: run the cell sweeps of kernels 2, 4 and 7 only on the row segments whose results can
     * change: a segment is left out when every cell its stencils reach ended the previous step
     * GAS with phi exactly 0 and the segment itself ended the step before so too (both buffers
     * of every pair then hold its results already). Honoured only on steady steps
     * (pending_stream 1, thermal_list_ready 1 or 2, arc_applied 0, forces_from_fields 0,
     * force_reset_all 0), which a caller may pass only when nothing was written between this
     * step and the waam_step_fused call before. 0: every segment is swept and the quiet record
     * restarts. The results are the same either way. */
    int32_t quiet_skip;
    /* arc */
    float arc_i, arc_j, arc_k_override, k_robot;
    float Q_w, eta, dt, dx, dx3, sigma, pen;
    float g_sgn, g_ff, g_fr, g_af, g_ar, g_b, g_c;
    /* thermal */
    float alpha_lu, cp_rho, L_rho, T_s, T_l, T_cap, T_amb, h_conv, eps_rad, sigma_sb, T800, T1100;
    /* forces */
    float gamma_lu, cos_t, sin_t, dgamma, p_sigma, p_pa, rho_phys, g_lu, beta, F_cap, u_cap;
    /* lattice Boltzmann */
    float omega_s, omega_b, C_darcy;
    /* freeze */
    float dir_x, dir_y, dir_z, lookback, T_freeze;
    /* recoil (recoil_pa and recoil_T_ref: constant form) and gas shear (peak stress, Pa) */
    float recoil_pa, recoil_T_ref, P_vap_ref, T_boil, L_vap, R_vap, C_acc, gas_tau;
    /* temperature-dependent material (reference coupled_step.py:168-245, 307-312, 364-373,
     * 426-449): 1 takes each FLUID/IFACE cell's cp, k, mu and dgamma/dT from `tables` at its
     * incoming temperature and `refs` on SOLID and GAS cells (refresh_thermal_properties), and
     * runs the _variable forms of conduction, floor, ceilings, phase map (with H_sol_tab,
     * H_liq_tab), Marangoni, remelt and trailing solidification; the four property fields of
     * waam_step_fields must be set. variable_tau 1 (with collision 2) collides with tau per
     * cell (collide_srt_variable_tau). Row-segment skipping is off on these steps. */
    int32_t use_tables, variable_tau;
    float H_sol_tab, H_liq_tab;
    waam_property_refs refs;
    waam_material_tables tables;
} waam_step_params;

int waam_step_fused(const waam_step_params *p, const waam_step_fields *f, void *stream);
/* A step on which a droplet detaches (reference solvers/coupled_step.py:97-166 followed by the
 * rest of the step), with everything the host would read back kept on the device: the previous
 * step's deferred stream is materialised; the torch column's surface height (with the robot
 * floor, as waam_step_fused resolves it) gives arc_k; arc heat and the ceiling are applied to H;
 * the forces are cleared; tracers are injected at the arc; the droplet is deposited with the
 * reference's retry loop (coupled_step.py:138-154: up to 10 sweeps, footprint + 1.5 cells up to
 * 16, until 98 % of the volume is placed) and its placed volume added to `ledger`; the momentum
 * impact and the impact pressure follow; then the fused step runs with that arc_k and those
 * forces. Populations: the deposition and the momentum impact work on fp32, so with 16-bit
 * storage f32_src/f32_dst (the fp32 pair of f_src/f_dst) are unpacked into and packed back
 * from; with fp32 storage they equal f_src/f_dst. No counterpart entry in the reference: its
 * droplet block reads the placed volume back on the host after every sweep. */
typedef struct waam_droplet {
    float *f32_src, *f32_dst;
    double drop_vol;        /* m^3 of metal to place (drop_mass / rho) */
    double footprint_r;     /* first footprint radius, cells */
    float arc_i, arc_j;     /* cells, as the step's params */
    float drop_r, T_drop, cp_rho, L_rho, rho0, cell_vol;
    float *vol_acc, *real_acc, *result; /* one float each, and two floats of sweep result */
    double *ledger;         /* [0] += placed volume (m^3); [1] += 1 when below 98 % of drop_vol */
    /* tracers (kernels.py:3186-3203): spawn at (tracer_x, tracer_y, (arc_k + 1) * dx) */
    float *tracer_pos;
    int32_t *tracer_active, *tracer_head;
    int32_t max_tracers, spawn;
    uint32_t seed;
    float tracer_x, tracer_y, sigma_m;
    double dx;
    /* impact (physics/weld_forces.py apply_droplet_impact): momentum (lattice velocity) and
     * pressure (Pa over the droplet's radius) */
    int32_t momentum, pressure;
    float vx, vy, vz, impact_pa, dt, dx_f, rho_ref;
} waam_droplet;
int waam_step_fused_droplet(const waam_step_params *p, const waam_step_fields *f,
                            const waam_droplet *drop, void *stream);
/* bytes of the quiet record (waam_step_fields.quiet) for a grid; 0 for invalid sizes */
int64_t waam_quiet_bytes(int nx, int ny, int nz, int layout);
/* Allocation-free variant of waam_step_fused (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same kernels and results as waam_step_fused; f->scratch, f->list_thermal
 * and f->list_collide are ignored, and a device array inside the library takes their place
 * (workspace plus two cell lists of waam_noalloc_max_cells() entries: 16.0 MiB + 0.7 MiB per
 * GPU, held from the first call on that GPU until the process ends). Grids larger than
 * waam_noalloc_max_cells() cells are rejected with WAAM_ERR_INVALID. The array keeps the last
 * call's contents: state that waam_step_fused carries in the workspace from one call to the next
 * (the thermal list when thermal_list_ready is set) is carried in it, so the first call after a
 * waam_step_fused call, or after a call on another grid, must pass thermal_list_ready 0. Every
 * call on one GPU shares its array: calls that overlap on one GPU (two streams, two threads)
 * race. Calls no hipMalloc, hipFree or allocating library routine. */
int waam_step_fused_noalloc(const waam_step_params *p, const waam_step_fields *f, void *stream);
int64_t waam_noalloc_max_cells(void);
/* Materialise a deferred stream: f_dst = pull(f_src) on cells not SOLID in sflags (the flags the
 * step that deferred it saw); populations in the WAAM_POP_* form pop_dtype. */
int waam_stream_pending(const void *f_src, void *f_dst, const uint8_t *sflags, int nx, int ny,
                        int nz, int layout, int pop_dtype, void *stream);
/* Convert n_cells x 19 populations between fp32 and the storage form pop_dtype (WAAM_POP_F16 or
 * WAAM_POP_BF16: f - w_q, rounded to nearest even; unpack adds w_q back in fp32). q-major, as
 * the population arrays are: element q * n_cells + c has weight w_q. Unpacking then packing
 * returns the same 16-bit words for every word that packing (or the fused collision's store)
 * produces. */
int waam_pack_populations(const float *src, void *dst, int64_t n_cells, int pop_dtype,
                          void *stream);
int waam_unpack_populations(const void *src, float *dst, int64_t n_cells, int pop_dtype,
                            void *stream);
/* The step's arc weight total, energy per unit weight and arc column height, as written into
 * the workspace by the last waam_step_fused (three device floats, in that order). Blocking. */
int waam_step_arc_readback(const void *scratch, float *host_out3);
/* The arc_k the last waam_step_fused_droplet used (resolved before its deposition). Blocking. */
int waam_step_droplet_arc_k(const void *scratch, float *host_out);

/* ---- captured steps: a run of n_steps consecutive waam_step_fused calls as one hipGraph.
 * Replaces n_steps calls of coupled_step (reference solvers/coupled_step.py:73) made by
 * WAAMTwin.step (twin.py:533-566) on steps with no droplet; the reference launches each kernel
 * from Python every step.
 * waam_graph_capture records the steps on a stream of its own (nothing runs): step k uses
 * f_even for even k and f_odd for odd k (f_odd must be f_even with every in/out pair swapped,
 * the state after one step), and p for everything but the arc, which each step reads from a
 * device track. p must describe a steady step (pending_stream 1, thermal_list_ready 1 or 2,
 * arc_applied 0, use_arc_k_override 0, force_reset_all 0, forces_from_fields 0): the captured
 * run must follow a fused step with nothing written in between. 1 <= n_steps <= 4096.
 * waam_graph_launch fills the track with the n_steps arc centres (cells, as the host computes
 * p.arc_i / p.arc_j; the columns are their truncations clamped to the grid), robot heights
 * k_robot (NULL: p.k_robot every step) and torch directions dir (3 per step; NULL: p.dir_*),
 * copies it on `stream` and launches the graph there.
 * It blocks only until the copy made two launches earlier has run. Launches of one graph
 * must use one stream; they run the same kernels, with the same results, as n_steps
 * waam_step_fused calls. Kernel timing (waam_timing_begin) does not see captured kernels.
 * waam_graph_destroy waits for the device to drain, then frees the graph. */
typedef struct waam_step_graph waam_step_graph;
int waam_graph_capture(const waam_step_params *p, const waam_step_fields *f_even,
                       const waam_step_fields *f_odd, int n_steps, waam_step_graph **out);
/* waam_graph_capture with a per-step plan (no counterpart in the reference): kinds[k] 0 is a
 * steady step, 1 a droplet step (waam_step_fused_droplet with drops[k]; its arc still comes
 * from the track), 2 the step after a droplet step (it clears every cell's forces). The
 * droplet parameters are recorded into the graph: a run with a droplet is captured for that
 * droplet. kinds NULL is waam_graph_capture. */
int waam_graph_capture_plan(const waam_step_params *p, const waam_step_fields *f_even,
                            const waam_step_fields *f_odd, int n_steps, const int32_t *kinds,
                            const waam_droplet *drops, waam_step_graph **out);
int waam_graph_launch(waam_step_graph *g, const double *arc_i, const double *arc_j,
                      const float *k_robot, const float *dir, void *stream);
int waam_graph_steps(const waam_step_graph *g);
int waam_graph_destroy(waam_step_graph *g);

/* ---- instrumentation (no counterpart in the reference) -------------------------------------
 * HIP-event timing of one kernel of waam_step_fused, recorded on the stream that kernel is
 * launched on: waam_timing_begin arms it for up to max_launches launches of kernel `which`
 * (one of WAAM_K_*); waam_timing_begin_every records only every `every`-th launch of it (the
 * events add a few microseconds around each recorded dispatch); waam_timing_end waits for the
 * recorded events, returns the summed duration and the number of timed launches, and disarms.
 * Process-wide, one kernel at a time; not for concurrent use from several threads. */
enum {
    WAAM_K_BEGIN = 0,
    WAAM_K_THERMAL = 1,
    WAAM_K_THERMAL_SURFACE = 2,
    WAAM_K_VOF_FLAGS = 3,
    WAAM_K_COLLIDE = 4,
    WAAM_K_TRACERS = 5,
    WAAM_K_STREAM = 6,
    WAAM_K_PULL = 7, /* the population pull when it runs ahead of kernel 7 in a launch of its own */
    WAAM_K_COUNT = 8
};
int waam_timing_begin(int which, int max_launches);
int waam_timing_begin_every(int which, int max_launches, int every);
int waam_timing_end(double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* WAAM_HIP_H */
