//! Rust binding of `libaqua.so` (include/aqua/aqua.h), the MI355X (gfx950)
//! path tracer and neural radiance cache behind the reference crate's own
//! data formats: the scene JSON (scenes/cbox.json, scenes/room.json), the
//! integrator JSON (scenes/integrator.json) and BSON `.mesh` geometry. The
//! reference crate `arukas` (Cargo.toml:1-30) is where this binding plugs in:
//! its host loads those files and calls `Renderer::render`.
//!
//! `ffi` mirrors the C ABI one to one (`#[repr(C)]` structs of fixed-width
//! scalars, `extern "C"` functions returning an `int` status); the safe
//! layer wraps the opaque handles with `Drop` and turns statuses into
//! `Result<_, Error>` carrying `aqua_last_error()`.
#![allow(non_camel_case_types)]

pub mod ffi {
    use std::os::raw::{c_char, c_int, c_void};

    pub const AQUA_ABI_VERSION: c_int = 1;

    pub const AQUA_OK: c_int = 0;
    pub const AQUA_ERR_IO: c_int = 1;
    pub const AQUA_ERR_PARSE: c_int = 2;
    pub const AQUA_ERR_MISSING_MESH: c_int = 3;
    pub const AQUA_ERR_NO_DEVICE: c_int = 4;
    pub const AQUA_ERR_OOM: c_int = 5;
    pub const AQUA_ERR_HIP: c_int = 6;
    pub const AQUA_ERR_INVALID_ARG: c_int = 7;
    pub const AQUA_ERR_NONFINITE: c_int = 8;
    pub const AQUA_ERR_UNSUPPORTED: c_int = 9;

    pub const AQUA_PROCGEN_BOX_ROOM: u32 = 1;
    pub const AQUA_PROCGEN_INTERIOR: u32 = 2;
    pub const AQUA_INTEGRATOR_PT: u32 = 0;
    pub const AQUA_INTEGRATOR_NRC: u32 = 1;
    pub const AQUA_SAMPLER_INDEPENDENT: u32 = 0;
    pub const AQUA_SAMPLER_STRATIFIED: u32 = 1;
    pub const AQUA_NRC_F16: u32 = 0;
    pub const AQUA_NRC_BF16: u32 = 1;
    pub const AQUA_NRC_FP8: u32 = 2;

    #[repr(C)]
    pub struct aqua_scene {
        _private: [u8; 0],
    }
    #[repr(C)]
    pub struct aqua_renderer {
        _private: [u8; 0],
    }
    #[repr(C)]
    pub struct aqua_nrc {
        _private: [u8; 0],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_load_opts {
        pub skip_missing: u32,
        pub image_as_constant: u32,
        pub emission_from_mtl: u32,
        pub reserved: [u32; 5],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_material {
        pub base_color: [f32; 3],
        pub metallic: f32,
        pub emission: [f32; 3],
        pub roughness: f32,
        pub specular: f32,
        pub base_color_tex: u32,
        pub pad: [f32; 2],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_camera {
        pub cam_to_world: [f32; 12],
        pub tan_half_fov: f32,
        pub aspect: f32,
        pub width: u32,
        pub height: u32,
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_point_light {
        pub pos: [f32; 3],
        pub intensity: [f32; 3],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_scene_info {
        pub n_triangles: u32,
        pub n_materials: u32,
        pub n_shapes: u32,
        pub n_skipped_shapes: u32,
        pub n_emissive_triangles: u32,
        pub n_point_lights: u32,
        pub n_image_textures: u32,
        pub n_vertices: u32,
        pub bounds_lo: [f32; 3],
        pub bounds_hi: [f32; 3],
        pub n_images: u32,
        pub has_texcoords: u32,
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_procgen_params {
        pub kind: u32,
        pub seed: u32,
        pub n_meshes: u32,
        pub target_tris: u32,
        pub width: u32,
        pub height: u32,
        pub reserved: [u32; 2],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_integrator_cfg {
        pub type_: u32,
        pub spp: u32,
        pub max_depth: u32,
        pub batch_size: u32,
        pub training_iters: u32,
        pub learning_rate: f32,
        pub visualize_cache: u32,
        pub seed: u32,
        pub rr_depth: u32,
        pub nrc_precision: u32,
        pub nrc_train_paths: u32,
        pub nrc_query_vertex: u32,
        pub sampler: u32,
        pub reserved: [u32; 3],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_frame_stats {
        pub camera_rays: u64,
        pub extension_rays: u64,
        pub shadow_rays: u64,
        pub nonfinite_samples: u64,
        pub ms_frame: f64,
        pub ms_intersect: f64,
        pub ms_shade: f64,
        pub ms_shadow: f64,
        pub n_intersect_launches: u32,
        pub n_passes: u32,
        pub nodes_visited: u64,
        pub tris_tested: u64,
        pub shadow_nodes_visited: u64,
        pub shadow_tris_tested: u64,
        pub nrc_queries: u32,
        pub nrc_train_records: u32,
        pub nrc_loss: f32,
        pub ms_nrc: f32,
        pub spill_node_visits: u64,
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_renderer_opts {
        pub samples_per_pass: u32,
        pub count_traversal: u32,
        pub timing: u32,
        pub sort_materials: u32,
        pub train_ahead: u32,
        pub reserved: [u32; 3],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_nrc_cfg {
        pub n_levels: u32,
        pub log2_table: u32,
        pub base_res: u32,
        pub level_scale: f32,
        pub learning_rate: f32,
        pub beta1: f32,
        pub beta2: f32,
        pub eps: f32,
        pub train_precision: u32,
        pub seed: u32,
        pub max_batch: u32,
        pub reserved: [u32; 5],
    }

    #[repr(C)]
    #[derive(Debug, Default, Clone, Copy)]
    pub struct aqua_nrc_stats {
        pub ms_total: f64,
        pub ms_fwd_bwd: f64,
        pub ms_wgrad: f64,
        pub ms_adam: f64,
        pub loss: f32,
        pub steps: u32,
        pub rows: u32,
        pub ms_scatter: f32,
        pub reserved: [u32; 4],
    }

    extern "C" {
        // ---- scene
        pub fn aqua_scene_load(scene_json_path: *const c_char, opts: *const aqua_load_opts, out: *mut *mut aqua_scene) -> c_int;
        pub fn aqua_scene_from_arrays(n_triangles: u32, verts: *const f32, mat_ids: *const u32, n_materials: u32,
                                      materials: *const aqua_material, camera: *const aqua_camera, n_point_lights: u32,
                                      point_lights: *const aqua_point_light, out: *mut *mut aqua_scene) -> c_int;
        pub fn aqua_scene_get_info(scene: *const aqua_scene, out: *mut aqua_scene_info) -> c_int;
        pub fn aqua_scene_get_triangles(scene: *const aqua_scene, verts: *mut f32, mat_ids: *mut u32) -> c_int;
        pub fn aqua_scene_get_materials(scene: *const aqua_scene, out: *mut aqua_material) -> c_int;
        pub fn aqua_scene_get_camera(scene: *const aqua_scene, out: *mut aqua_camera) -> c_int;
        pub fn aqua_scene_get_point_lights(scene: *const aqua_scene, out: *mut aqua_point_light) -> c_int;
        pub fn aqua_scene_get_texcoords(scene: *const aqua_scene, uv: *mut f32) -> c_int;
        pub fn aqua_scene_get_image(scene: *const aqua_scene, k: u32, width: *mut u32, height: *mut u32, rgba: *mut u8) -> c_int;
        pub fn aqua_scene_set_texcoords(scene: *mut aqua_scene, uv: *const f32) -> c_int;
        pub fn aqua_scene_add_image(scene: *mut aqua_scene, width: u32, height: u32, rgba: *const u8, index: *mut u32) -> c_int;
        pub fn aqua_scene_destroy(scene: *mut aqua_scene);
        // ---- images
        pub fn aqua_image_load(path: *const c_char, width: *mut u32, height: *mut u32, rgba: *mut u8) -> c_int;
        pub fn aqua_image_write_png(path: *const c_char, width: u32, height: u32, channels: u32, pixels: *const u8) -> c_int;
        pub fn aqua_image_write_pfm(path: *const c_char, width: u32, height: u32, rgb: *const f32) -> c_int;
        // ---- procedural scenes
        pub fn aqua_procgen_write(params: *const aqua_procgen_params, out_dir: *const c_char, scene_path: *mut c_char, cap: usize) -> c_int;
        // ---- integrator
        pub fn aqua_integrator_default(out: *mut aqua_integrator_cfg);
        pub fn aqua_integrator_load(integrator_json_path: *const c_char, out: *mut aqua_integrator_cfg) -> c_int;
        // ---- renderer
        pub fn aqua_renderer_create(scene: *const aqua_scene, cfg: *const aqua_integrator_cfg, device: c_int, out: *mut *mut aqua_renderer) -> c_int;
        pub fn aqua_renderer_set_opts(r: *mut aqua_renderer, opts: *const aqua_renderer_opts) -> c_int;
        pub fn aqua_render_frame(r: *mut aqua_renderer, camera: *const aqua_camera, frame: u32, spp: u32, stats: *mut aqua_frame_stats) -> c_int;
        pub fn aqua_film_read(r: *mut aqua_renderer, rgb: *mut f32, spp: *mut u32) -> c_int;
        pub fn aqua_film_tonemap(r: *mut aqua_renderer, exposure: f32, rgba: *mut u8) -> c_int;
        pub fn aqua_render_oracle(r: *mut aqua_renderer, camera: *const aqua_camera, frame: u32, spp_begin: u32, spp_count: u32) -> c_int;
        pub fn aqua_oracle_read(r: *mut aqua_renderer, rgb: *mut f64, spp: *mut u32) -> c_int;
        pub fn aqua_film_error(r: *mut aqua_renderer, relmse: *mut f64, relrmse: *mut f64) -> c_int;
        pub fn aqua_primary_hits(r: *mut aqua_renderer, camera: *const aqua_camera, frame: u32, sample: u32, prim_ids: *mut u32, t_hit: *mut f32) -> c_int;
        pub fn aqua_trace_closest(r: *mut aqua_renderer, n: u32, rays: *const f32, prim_ids: *mut u32, t_hit: *mut f32) -> c_int;
        pub fn aqua_trace_closest_noalloc(r: *mut aqua_renderer, n: u32, d_rays: *const f32, d_prim_ids: *mut u32, d_t_hit: *mut f32) -> c_int;
        pub fn aqua_trace_occluded(r: *mut aqua_renderer, n: u32, rays: *const f32, occluded: *mut u8) -> c_int;
        pub fn aqua_renderer_destroy(r: *mut aqua_renderer);
        // ---- neural radiance cache
        pub fn aqua_nrc_default_cfg(out: *mut aqua_nrc_cfg);
        pub fn aqua_nrc_create(device: c_int, cfg: *const aqua_nrc_cfg, out: *mut *mut aqua_nrc) -> c_int;
        pub fn aqua_nrc_destroy(nrc: *mut aqua_nrc);
        pub fn aqua_nrc_param_counts(nrc: *const aqua_nrc, n_table: *mut u64, n_mlp: *mut u32) -> c_int;
        pub fn aqua_nrc_get_params(nrc: *mut aqua_nrc, table: *mut f32, mlp: *mut f32) -> c_int;
        pub fn aqua_nrc_set_params(nrc: *mut aqua_nrc, table: *const f32, mlp: *const f32) -> c_int;
        pub fn aqua_nrc_fp8_scales(nrc: *mut aqua_nrc, wscale: *mut f32, ascale: *mut f32) -> c_int;
        pub fn aqua_nrc_query_order(nrc: *mut aqua_nrc, n: u32, records: *const f32, order: *mut u32) -> c_int;
        pub fn aqua_nrc_checkpoint(nrc: *mut aqua_nrc, restore: c_int) -> c_int;
        pub fn aqua_nrc_infer(nrc: *mut aqua_nrc, n: u32, records: *const f32, precision: u32, rgb: *mut f32) -> c_int;
        pub fn aqua_nrc_train(nrc: *mut aqua_nrc, n: u32, records: *const f32, targets: *const f32, steps: u32, stats: *mut aqua_nrc_stats) -> c_int;
        pub fn aqua_nrc_gradients(nrc: *mut aqua_nrc, n: u32, records: *const f32, targets: *const f32, g_table: *mut f32, g_mlp: *mut f32, loss: *mut f32) -> c_int;
        pub fn aqua_nrc_train_device(nrc: *mut aqua_nrc, n: u32, d_records: *const c_void, d_targets: *const c_void, steps: u32, stats: *mut aqua_nrc_stats) -> c_int;
        pub fn aqua_renderer_nrc(r: *mut aqua_renderer) -> *mut aqua_nrc;
        pub fn aqua_renderer_set_nrc_enabled(r: *mut aqua_renderer, enabled: c_int) -> c_int;
        pub fn aqua_renderer_set_nrc_precision(r: *mut aqua_renderer, precision: u32) -> c_int;
        pub fn aqua_renderer_nrc_checkpoint(r: *mut aqua_renderer, restore: c_int) -> c_int;
        pub fn aqua_renderer_nrc_batch(r: *mut aqua_renderer, n: *mut u32, d_records: *mut *const c_void, d_targets: *mut *const c_void) -> c_int;
        // ---- misc
        pub fn aqua_camera_flythrough(scene: *const aqua_scene, seed: u32, frame: u32, n_frames: u32, fov_deg: f32, width: u32, height: u32, out: *mut aqua_camera) -> c_int;
        pub fn aqua_last_error() -> *const c_char;
        pub fn aqua_abi_version() -> c_int;
        pub fn aqua_device_count(count: *mut c_int) -> c_int;
        pub fn aqua_scene_bvh(scene: *const aqua_scene, n_nodes: *mut u32, nodes: *mut f32, order: *mut u32, max_depth: *mut u32) -> c_int;
        pub fn aqua_scene_bvh_device(scene: *const aqua_scene, n_nodes: *mut u32, words: *mut u32) -> c_int;
        pub fn aqua_fp_probe(n: u32, a: *const f32, b: *const f32, c: *const f32, u: *const f32, out: *mut f32) -> c_int;
    }
}

use std::ffi::{CStr, CString};
use std::fmt;
use std::os::raw::c_int;
use std::path::Path;
use std::ptr;

/// A non-zero status of the C ABI with the library's message.
#[derive(Debug, Clone)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for Error {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "aqua error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

pub type Result<T> = std::result::Result<T, Error>;

fn check(code: c_int) -> Result<()> {
    if code == ffi::AQUA_OK {
        return Ok(());
    }
    let message = unsafe {
        let p = ffi::aqua_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    };
    Err(Error { code, message })
}

fn c_path(p: &Path) -> Result<CString> {
    CString::new(p.to_string_lossy().as_bytes()).map_err(|_| Error {
        code: ffi::AQUA_ERR_INVALID_ARG,
        message: format!("path {:?} holds a NUL byte", p),
    })
}

/// Options of `Scene::load` (aqua_load_opts).
#[derive(Debug, Default, Clone, Copy)]
pub struct LoadOptions {
    pub skip_missing: bool,
    pub image_as_constant: bool,
    pub emission_from_mtl: bool,
}

/// A loaded scene: the flattened triangles, materials, camera, point lights
/// and image textures of a reference-schema scene JSON.
pub struct Scene {
    raw: *mut ffi::aqua_scene,
}

impl Scene {
    pub fn load<P: AsRef<Path>>(path: P, opts: LoadOptions) -> Result<Scene> {
        let p = c_path(path.as_ref())?;
        let o = ffi::aqua_load_opts {
            skip_missing: opts.skip_missing as u32,
            image_as_constant: opts.image_as_constant as u32,
            emission_from_mtl: opts.emission_from_mtl as u32,
            reserved: [0; 5],
        };
        let mut raw = ptr::null_mut();
        check(unsafe { ffi::aqua_scene_load(p.as_ptr(), &o, &mut raw) })?;
        Ok(Scene { raw })
    }

    pub fn info(&self) -> Result<ffi::aqua_scene_info> {
        let mut i = ffi::aqua_scene_info::default();
        check(unsafe { ffi::aqua_scene_get_info(self.raw, &mut i) })?;
        Ok(i)
    }

    pub fn camera(&self) -> Result<ffi::aqua_camera> {
        let mut c = ffi::aqua_camera::default();
        check(unsafe { ffi::aqua_scene_get_camera(self.raw, &mut c) })?;
        Ok(c)
    }

    pub fn as_ptr(&self) -> *const ffi::aqua_scene {
        self.raw
    }
}

impl Drop for Scene {
    fn drop(&mut self) {
        unsafe { ffi::aqua_scene_destroy(self.raw) }
    }
}

/// The integrator JSON (scenes/integrator.json:2-8).
pub fn integrator_load<P: AsRef<Path>>(path: P) -> Result<ffi::aqua_integrator_cfg> {
    let p = c_path(path.as_ref())?;
    let mut c = ffi::aqua_integrator_cfg::default();
    check(unsafe { ffi::aqua_integrator_load(p.as_ptr(), &mut c) })?;
    Ok(c)
}

pub fn integrator_default() -> ffi::aqua_integrator_cfg {
    let mut c = ffi::aqua_integrator_cfg::default();
    unsafe { ffi::aqua_integrator_default(&mut c) };
    c
}

/// The renderer on one GPU: the scene resident in HBM, path queues, films
/// and (for "type": "nrc") the neural radiance cache.
pub struct Renderer {
    raw: *mut ffi::aqua_renderer,
    width: u32,
    height: u32,
}

impl Renderer {
    pub fn new(scene: &Scene, cfg: &ffi::aqua_integrator_cfg, device: i32) -> Result<Renderer> {
        let mut raw = ptr::null_mut();
        check(unsafe { ffi::aqua_renderer_create(scene.raw, cfg, device, &mut raw) })?;
        let cam = scene.camera()?;
        Ok(Renderer { raw, width: cam.width, height: cam.height })
    }

    /// Renders `spp` samples per pixel of `frame`; `camera` None = the scene's.
    pub fn render(&mut self, camera: Option<&ffi::aqua_camera>, frame: u32, spp: u32) -> Result<ffi::aqua_frame_stats> {
        let mut st = ffi::aqua_frame_stats::default();
        let cam = camera.map_or(ptr::null(), |c| c as *const _);
        check(unsafe { ffi::aqua_render_frame(self.raw, cam, frame, spp, &mut st) })?;
        if let Some(c) = camera {
            self.width = c.width;
            self.height = c.height;
        }
        Ok(st)
    }

    /// Mean radiance of the last frame, RGB per pixel, row 0 = top.
    pub fn film(&mut self) -> Result<Vec<f32>> {
        let mut rgb = vec![0f32; 3 * self.width as usize * self.height as usize];
        let mut spp = 0u32;
        check(unsafe { ffi::aqua_film_read(self.raw, rgb.as_mut_ptr(), &mut spp) })?;
        Ok(rgb)
    }

    /// The last frame as 8-bit sRGB RGBA, x 2^exposure (computed on the GPU).
    pub fn tonemap(&mut self, exposure: f32) -> Result<Vec<u8>> {
        let mut px = vec![0u8; 4 * self.width as usize * self.height as usize];
        check(unsafe { ffi::aqua_film_tonemap(self.raw, exposure, px.as_mut_ptr()) })?;
        Ok(px)
    }

    pub fn size(&self) -> (u32, u32) {
        (self.width, self.height)
    }

    pub fn as_mut_ptr(&mut self) -> *mut ffi::aqua_renderer {
        self.raw
    }
}

impl Drop for Renderer {
    fn drop(&mut self) {
        unsafe { ffi::aqua_renderer_destroy(self.raw) }
    }
}

/// Writes 8-bit RGBA pixels (row 0 = top) as a PNG file.
pub fn write_png<P: AsRef<Path>>(path: P, width: u32, height: u32, rgba: &[u8]) -> Result<()> {
    if rgba.len() != 4 * width as usize * height as usize {
        return Err(Error { code: ffi::AQUA_ERR_INVALID_ARG, message: "pixel count does not match the size".into() });
    }
    let p = c_path(path.as_ref())?;
    check(unsafe { ffi::aqua_image_write_png(p.as_ptr(), width, height, 4, rgba.as_ptr()) })
}

/// Number of HIP devices visible (0 without a GPU).
pub fn device_count() -> Result<i32> {
    let mut n: c_int = 0;
    check(unsafe { ffi::aqua_device_count(&mut n) })?;
    Ok(n)
}
