/*
 * silt_core.h -- the two-layer FluidCore of libsilt_amd.so.
 *
 * Replaces FluidCore (src/fluid_core.rs:4808-4955) on the sheet path: the
 * 40 m coarse window (200 x 200 cells of 0.2 m) and the 2 cm patch
 * (160 x 160 cells) laid under the tires each step.  State lives on the GPU.
 * Closures of the reference API (map_water_bed, lay_slurry_mask,
 * read_substrate) become arrays sampled at coarse cell centres
 * origin + (i + 0.5) * 0.2, row-major idx = j * 200 + i.
 * A handle is not thread-safe; calls on one handle are serialised on its own
 * HIP stream.  Negative returns are SILT_E* codes (silt_sheet.h).
 */
#ifndef SILT_CORE_H
#define SILT_CORE_H

#include <stddef.h>
#include <stdint.h>

#include "silt_rut.h"
#include "silt_sheet.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SILT_WINDOW_RES 200
#define SILT_FINE_RES 160

typedef struct silt_core silt_core;

/* TirePush, src/fluid_core.rs:357-368. */
typedef struct silt_tire_push {
    float x, z, forward_x, forward_z, radius, width, speed, slip;
} silt_tire_push;

/* SprayEvent, src/fluid_core.rs:371-382 (kind 0 = drop). */
typedef struct silt_spray_event {
    float x, y, z, vx, vy, vz, volume;
    int32_t kind;
} silt_spray_event;

/* FluidLedger (src/fluid_core.rs:483-509) restricted to the sheet path, FP64. */
typedef struct silt_fluid_ledger {
    double water_m3, slurry_m3, frozen_water_m3, frozen_slurry_m3, spray_m3, total_m3;
} silt_fluid_ledger;

/* FluidLedger (src/fluid_core.rs:483-509) with all 15 fields, FP64.  The
 * suspended-solid lines are the wheel-lifted column (silt_core_get_suspended),
 * what settled back, what is parked off the window and the entrained source;
 * total_m3 adds the first three to the sheet's total, as the reference does.
 * The dust, fog and river-reach lines belong to subsystems this library does
 * not build (the air grid, river reaches) and are 0. */
typedef struct silt_fluid_ledger_full {
    double water_m3, slurry_m3, dust_air_m3, dust_settled_m3, frozen_water_m3, frozen_slurry_m3, frozen_dust_m3,
        spray_m3, fog_air_m3, suspended_m3, settled_suspended_m3, frozen_suspended_m3, entrained_m3, reach_m3,
        total_m3;
} silt_fluid_ledger_full;

/* CoarseCell (src/fluid_core.rs:450-463); turbidity = (suspended + stirred
 * solids) / water, clamped to [0, 1]; settled dust is 0 (no air grid). */
typedef struct silt_coarse_cell {
    float x, z, bed_m, water_m, slurry_m, moisture, u, v, settled_dust_m, turbidity;
} silt_coarse_cell;

/* GroundSample (src/fluid_core.rs:440-448). */
typedef struct silt_ground_sample {
    float water_m, slurry_m, u, v, settled_dust_m, moisture, bed_m;
} silt_ground_sample;

/* FluidCore::new(center) (src/fluid_core.rs:4958-5060).  path: SILT_PATH_*
 * for the sheet steps.  NULL on failure. */
silt_core *silt_core_create(float center_x, float center_z, int path);
void silt_core_destroy(silt_core *c);

/* set_fine_pace / set_fine_sub_cap (src/fluid_core.rs:5072-5091),
 * set_near_field (src/fluid_core/near.rs:332-338).  coarse_confine picks the
 * coarse sheets' confinement: SILT_CONFINE_REFERENCE (in-place sweep, as the
 * reference steps a layer beside the other) or SILT_CONFINE_JACOBI. */
int silt_core_set_fine_pace(silt_core *c, int pace);
int silt_core_set_fine_sub_cap(silt_core *c, float cap);
int silt_core_set_near_field(silt_core *c, int on);
int silt_core_set_coarse_confine(silt_core *c, int confine);

/* window_bounds origin (src/fluid_core.rs:6977-6990): out2 = corner x, z. */
int silt_core_window_origin(const silt_core *c, float *out2);

/* map_water_bed (src/fluid_core.rs:5329-5355): bed and water per cell. */
int silt_core_map_water_bed(silt_core *c, const float *bed, const float *water);
/* pour_water / pour_slurry (src/fluid_core.rs:5424-5440 via paint :5820-5855).
 * water != 0 pours water; otherwise slurry of the grain code (Grain) and moisture. */
int silt_core_pour(silt_core *c, float x0, float x1, float z0, float z1, float depth, int water, int grain,
                   float moisture);
/* lay_slurry_mask (src/fluid_core.rs:5857-5880). */
int silt_core_lay_slurry(silt_core *c, const float *depth, int grain, float moisture);
/* read_substrate_grid (src/fluid_core.rs:5928-5962): bed, moisture, grain per cell. */
int silt_core_read_substrate(silt_core *c, const float *bed, const float *moisture, const uint8_t *grain);
/* push_tire (src/fluid_core.rs:5964-5966); at most 64 per step. */
int silt_core_push_tire(silt_core *c, const silt_tire_push *push);

/* FluidCore::step (src/fluid_core.rs:6034-6128) on the sheet path.  Up to
 * cap spray events are copied to events; *nevents gets the count the step
 * produced (may exceed cap).  Returns SILT_OK. */
int silt_core_step(silt_core *c, float dt, silt_spray_event *events, size_t cap, size_t *nevents);

/* ledger() (src/fluid_core.rs:6821-6883) for the sheets, parked fine volume and spray. */
int silt_core_ledger(silt_core *c, silt_fluid_ledger *out);

/* The FP64 lines that close the ledger, per layer ([0] water, [1] slurry).
 * total: the layer's volume with its frozen part (and, for water, the
 * spray).  cut: depth the sheet steps' dry cuts deleted.  residue: what the
 * FP32 transports and the surrogate's fluxes created or removed beyond the
 * exchanged fluxes (rounding and the max(0) clamp).  handoff: what the
 * window <-> 2 cm patch handoffs changed (the FP32 restrict sums, the near
 * field's lay-down).  source: what entered from the bed (the scour beside
 * tire contacts; slurry only).  base: the total at the last caller edit
 * (set_planes, pours, map_water_bed, read_substrate, lay_slurry*,
 * drain_over_bed, recenter) less the lines booked by then.  gap = total -
 * base - (residue - cut + handoff + source): volume no line accounts for.
 * The reference's ledger() has no such lines (FluidLedger,
 * src/fluid_core.rs:483-509). */
typedef struct silt_core_lines {
    double base_m3[2], cut_m3[2], residue_m3[2], handoff_m3[2], source_m3[2], total_m3[2], gap_m3[2];
} silt_core_lines;
int silt_core_ledger_lines(silt_core *c, silt_core_lines *out);

/* HydroSolid (src/fluid_core.rs:219-238): a box (the belly) or a tire
 * contact that the sheet treats as a wall wherever the free surface stands
 * above bottom_y.  yaw 0 points the long axis along +z.  contact: nonzero
 * for a tire contact (scour reads the shear beside it); radius > 0 makes the
 * buoyancy a submerged cylinder segment of that radius and width. */
typedef struct silt_hydro_solid {
    float x, z, yaw, half_length, half_width, bottom_y, top_y;
    int32_t contact;
    float radius, width;
} silt_hydro_solid;

/* HydroLoad (src/fluid_core.rs:241-253). */
typedef struct silt_hydro_load {
    float force_x, force_z, buoyancy_n, hull_buoyancy_n, wheel_buoyancy_n, displaced_m3, eroded_m3, hole_m;
    uint32_t wheel_blocks, belly_blocks;
} silt_hydro_load;

/* set_hydro_solids (src/fluid_core.rs:5257-5260): the walls for the next
 * steps (at most 64; n = 0 is open water).  With walls every layer step of
 * FluidCore::step is the walled sheet (step_pair :7954-8125: the grid's wall
 * mask, the body displacement, silt_step_sheet_walled's substeps), and the
 * window's step ends with scour_contacts (:6076-6078, :7364-7421). */
int silt_core_set_hydro_solids(silt_core *c, const silt_hydro_solid *solids, size_t n);

/* hydro_load (src/fluid_core.rs:7164-7249): pressure force on the walls,
 * buoyancy and the scour account, read from the window. */
int silt_core_hydro_load(silt_core *c, silt_hydro_load *out);

/* The foam and stirred-silt tracers (src/fluid_core.rs:6210-6750): every
 * silt_core_step ends with step_tracers (:6079), which carries foam
 * (semi-Lagrangian, 1.8 s life) and stirred silt (donor-cell, settling over
 * 4 s) on the window and spawns foam where tires cut the surface, at shores
 * and where the film runs supercritical, and silt where a tire slips and at
 * shores.  foam_crest's foam channel publishes the foam.
 * note_surface_foam (:6526-6538): the crest foam of a seated sheet, for the
 * caller that seats the rut water after the step (ThroughDrive). */
int silt_core_note_surface_foam(silt_core *c, float dt);
/* pond_ruts (src/fluid_core.rs:5509-5611): water in a tire cut (bed more
 * than 12 mm below the pour plane) rises to the mean free surface of the
 * still pond cells within 10 cells, and those cells give the same cubic
 * metres in proportion to their depth above 12 mm.  moved (may be NULL: then
 * the call does not wait for the GPU) = the cubic metres moved.  The FP64
 * change of the window's water, the f32 rounding of the move, is booked as
 * the water's handoff line (silt_core_ledger_lines). */
int silt_core_pond_ruts(silt_core *c, float *moved);
/* seat_rut_water (src/fluid_core.rs:5622-5736) with shove_rut_slurry
 * (:5740-5792).  The classify closure becomes two host arrays sampled at the
 * coarse cell centres (200 x 200): the pressed ground (bed + placed, m;
 * non-finite reads as 0) and the cut flags.  The pond surface is the median
 * free surface of the floor cells; cut and bank columns above it are capped,
 * the water they drop goes back onto the floor in proportion to depth, and
 * slurry scraped off a cut goes to the nearest ring of uncut cells.
 * has_floor = 1 and *eta = the pond surface when the window has at least 8
 * floor cells, else 0 and NaN (the reference's None).  eta / has_floor may be
 * NULL (no wait).  The FP64 changes of water and slurry are booked as
 * handoff lines. */
int silt_core_seat_rut_water(silt_core *c, const float *pressed, const uint8_t *is_rut, float *eta, int *has_floor);
/* seat_rut_sheet (src/world.rs:5539-5561): seat_rut_water classified from
 * the core's ground (silt_core_set_ground, cut by silt_core_stamp_ruts):
 * pressed = bed + placed, a cut where the ground lies more than 12 mm below
 * its initial height.  A window whose coarse sheet is dry and which holds no
 * cut is skipped.  seated (may be NULL: no wait) = 1 when a floor was found.
 * SILT_EINVAL without a ground. */
int silt_core_seat_rut_sheet(silt_core *c, int *seated);
/* The wheel-lifted suspended solids (src/fluid_core.rs:8750-9003): every
 * silt_core_step with pushes first lifts solids into the water column behind
 * each tire moving at 0.3 m/s or more (entrain_wake :8758-8808), and every
 * step lets them fall out at 4 mm/s and carries them with the water's v
 * (step_suspended :8750-8756, advect_suspended :8810-9003; the window is
 * never surface driven here, so no u and no eddy mixing).  They feed the
 * contract's turbidity with the stirred silt.  susp: 200 x 200 (m) or NULL;
 * accounts2 = FP64 [entrained from the bed, settled back] (m^3) or NULL. */
int silt_core_get_suspended(silt_core *c, float *susp, double *accounts2);
int silt_core_set_suspended(silt_core *c, const float *susp);
/* The window's grain plane out (Grain codes, src/soil.rs; 200 x 200 bytes). */
int silt_core_get_grain(silt_core *c, uint8_t *grain);
/* foam fraction and stirred silt (m), 200 x 200 each (either may be NULL);
 * accounts2 = FP64 silt volume stirred in and settled out (m^3), or NULL. */
int silt_core_get_tracers(silt_core *c, float *foam, float *turbid, double *accounts2);
/* Lay foam and silt planes (either may be NULL); cells with foam over 1e-4
 * become live foam. */
int silt_core_set_tracers(silt_core *c, const float *foam, const float *turbid);

/* The wall mask solid_mask (src/fluid_core.rs:7288-7362) lays from the
 * current planes: 200 x 200 bytes for the window (fine 0) or 160 x 160 for
 * the 2 cm patch while it is on (fine 1); 2 wall, 1 contact, 0 open. */
int silt_core_solid_mask(silt_core *c, int fine, uint8_t *out);

/* Coarse planes in and out, 200 x 200 each (any pointer may be NULL to skip). */
int silt_core_get_planes(silt_core *c, float *water, float *slurry, float *uw, float *vw, float *us, float *vs,
                         float *bed, float *moisture);
int silt_core_set_planes(silt_core *c, const float *water, const float *slurry, const float *uw, const float *vw,
                         const float *us, const float *vs, const float *bed, const float *moisture);

/* publish maps 1/3/4 of the render contract (src/render_contract.rs:424-496,
 * docs/RENDER_CONTRACT.md:35-38) on the 200 x 200 window: film rg32f
 * (water, slurry), flow rgba32f (u, v, speed, turbidity: coarse_cell's, 0
 * on a dry cell), ground_rut rgba32f (bed + placed, rut, bed, placed).
 * soil_* (host, 200 x 200) may be NULL:
 * then the core's ground (silt_core_set_ground) is published, or, with none
 * set, the sheet bed with no placed soil and no rut.  Host outputs. */
int silt_core_publish(silt_core *c, const float *soil_bed, const float *soil_placed, const float *soil_rut,
                      float *ground_rut4, float *film2, float *flow4);
/* The render contract's moisture (r32f), wet_edge (r32f) and foam_crest
 * (rg32f) maps on the 200 x 200 window (publish_with_snow,
 * src/render_contract.rs:457-500 and :526-569): moisture = max(sheet
 * moisture, soil wetness); wet_edge = 1 where a 4-neighbour inside the
 * window differs in is_wet (moisture >= 0.45 or film >= 0.004 m,
 * :261-263); foam_crest = (foam, placed + max(film - mean neighbour film,
 * 0)), foam from the foam tracer (silt_core_get_tracers).  Host
 * arrays; soil_wetness / soil_placed may be NULL (no soil texel; placed then
 * comes from silt_core_set_ground's ground, or is 0); any output may be NULL. */
int silt_core_publish_surface(silt_core *c, const float *soil_wetness, const float *soil_placed, float *moisture1,
                              float *wet_edge1, float *foam_crest2);
/* silt_core_publish and silt_core_publish_surface into device memory: every
 * pointer is a device pointer (the renderer's own buffers, which
 * ContractGpu::upload fills from host copies in the reference,
 * src/render_contract.rs:1146-1166), nothing goes through the host.
 * Asynchronous: the maps are written after the handle's pending work and
 * after what hip_stream holds when the call is made, and work enqueued on
 * hip_stream afterwards sees them.  Inputs and outputs may be NULL as in the
 * host calls. */
int silt_core_publish_device(silt_core *c, const float *soil_bed, const float *soil_placed, const float *soil_rut,
                             float *ground_rut4, float *film2, float *flow4, void *hip_stream);
int silt_core_publish_surface_device(silt_core *c, const float *soil_wetness, const float *soil_placed,
                                     float *moisture1, float *wet_edge1, float *foam_crest2, void *hip_stream);

/* The ground under the sheet on the window grid (SoilField bed and placed
 * soil sampled at the cell centres).  The initial ground is this bed; the
 * sheet bed becomes bed + placed (pull_substrate, src/world.rs:5511-5537).
 * placed may be NULL (no placed soil). */
int silt_core_set_ground(silt_core *c, const float *bed, const float *placed);
/* Bekker ruts of the contacts (at most 64) cut into that ground
 * (include/silt_rut.h), then the sheet bed is pulled from it again. */
int silt_core_stamp_ruts(silt_core *c, const silt_tire_contact *contacts, size_t n, float dt);
/* The ground planes out: bed, placed, rut = max(initial - bed, 0). */
int silt_core_get_ground(silt_core *c, float *bed, float *placed, float *rut);

/* fine_water_debug (src/fluid_core/near.rs:341-346): lay the 2 cm patch
 * under n pushes with dt = 0, copy its water (and, when not NULL, its u and v;
 * 160 x 160 each) and its world origin (out2), then restrict it again. */
int silt_core_fine_water_debug(silt_core *c, const silt_tire_push *pushes, size_t n, float *water, float *uw,
                               float *vw, float *origin2);

/* map_water_velocity (src/fluid_core.rs:5358-5372): water (u, v) per cell
 * where the water is deeper than 1e-6 m; non-finite values become 0. */
int silt_core_map_water_velocity(silt_core *c, const float *u, const float *v);
/* non_finite_count (src/fluid_core.rs:5374-5395) over the planes this
 * library holds (water, slurry, their velocities, bed, moisture). */
long silt_core_non_finite_count(silt_core *c);
/* recenter (src/fluid_core.rs:5409-5422, shift_coarse :9004-9226): move the
 * window to centre (x, z) snapped to the coarse grid.  Cells that slide off
 * are parked by world cell (water, slurry, velocities, moisture; when they
 * hold anything), cells that slide on are restored from the parked set or
 * start dry at the ambient moisture; bed and grain stay on their array index
 * (read the substrate again).  The ledger's frozen lines include the parked
 * volume.  silt_core_parked_cells: cells in the parked set. */
int silt_core_recenter(silt_core *c, float x, float z);
long silt_core_parked_cells(const silt_core *c);
/* window_bounds (src/fluid_core.rs:6977-6980): out3 = corner x, z, extent. */
int silt_core_window_bounds(const silt_core *c, float *out3);
/* lay_slurry_on_water (src/fluid_core.rs:5236-5254): slurry `depth` deep,
 * of grain (0 clay .. 4 other), on every cell with more than 12 mm of water;
 * moisture raised to 0.9.  Returns the volume laid, m^3 (f32 sum). */
float silt_core_lay_slurry_on_water(silt_core *c, float depth, int grain);
/* set_center (src/fluid_core.rs:5403-5406): the chase point, the window does
 * not move; silt_core_center reads it back. */
int silt_core_set_center(silt_core *c, float x, float z);
int silt_core_center(const silt_core *c, float *out2);
/* coarse_cell (src/fluid_core.rs:6998-7034): 1 and *out set, or 0 off the
 * window (the reference's None). */
int silt_core_coarse_cell(silt_core *c, size_t i, size_t j, silt_coarse_cell *out);
/* sample (src/fluid_core.rs:6806-6813, sample_coarse :9428-9454): the coarse
 * cell under (x, z) (its slurry velocity, as the reference returns); zeros
 * off the window (no parked cells here). */
int silt_core_sample(silt_core *c, float x, float z, silt_ground_sample *out);
/* ledger() with every FluidLedger field. */
int silt_core_ledger_full(silt_core *c, silt_fluid_ledger_full *out);
/* drain_over_bed (src/fluid_core.rs:5123-5214): up to `steps` gravity steps
 * of both coarse layers (no confinement, dry banks, pace 2), velocities damped
 * by exp(-0.55 dt), volume wiped by the dry cut put back on the deepest cell,
 * stopping 24 calm steps after the sheet has moved.  Returns the water volume
 * in m^3 (NaN on a device error); volumes are the reference's f32 sums. */
float silt_core_drain_over_bed(silt_core *c, float dt, unsigned steps);

/* Near-field residual tiles alive (FluidCore::near_tiles, near.rs:316-318). */
long silt_core_near_tiles(silt_core *c);

/* How a step treats the tires.  No counterpart in the reference, which always
 * runs the nested patch (SILT_TIRE_NESTED, the default and the reference's
 * step).  SILT_TIRE_COARSE: no patch, every tire shoves the coarse sheet (the
 * coarse step without a sub-grid model).  SILT_TIRE_SURROGATE: the coarse
 * step plus the tire sub-grid network's edge fluxes (include/silt_mlp.h) over
 * the cells the patch would cover, grown by 2; needs silt_core_set_surrogate.
 * For callers that trade the patch's detail for its cost. */
#define SILT_TIRE_NESTED 0
#define SILT_TIRE_COARSE 1
#define SILT_TIRE_SURROGATE 2
int silt_core_set_tire_model(silt_core *c, int model);
/* The network's parameters (SILT_MLP_PARAMS floats, silt_mlp.h order) and its
 * judge: every judge_every steps with tires (0: never) the nested step also
 * runs from the same state and its result is kept; when the two depths differ
 * by more than tol metres anywhere in the window, the next judge_every steps
 * run nested (fallback), and the step after them is judged again. */
int silt_core_set_surrogate(silt_core *c, const float *params, int judge_every, float tol);

typedef struct silt_surrogate_stats {
    float last_tire_ms;     /* GPU time of the last step's tire block: the 2 cm patch
                               (refine, fine step, restrict) or the network and its fluxes */
    float last_judge_err_m; /* largest depth difference at the last judged step; -1 before one */
    float last_judge_rms_m; /* root mean square depth difference over the network's region, both
                               layers, at the last judged step; -1 before one */
    int judged;             /* judged steps so far */
    int fallbacks;          /* judged steps over the tolerance */
    int last_mode;          /* SILT_TIRE_* whose result the last step kept */
} silt_surrogate_stats;
int silt_core_surrogate_stats(const silt_core *c, silt_surrogate_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* SILT_CORE_H */
