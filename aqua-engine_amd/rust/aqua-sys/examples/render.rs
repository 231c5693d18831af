// The renderer entry point a host of the reference would run:
//     cargo run --example render -- scenes/cbox.json scenes/integrator.json out.png
// Loads the scene JSON (its meshes and textures) and the integrator JSON,
// renders one frame on GPU 0 and writes it as a PNG.
use aqua_sys::{integrator_load, write_png, LoadOptions, Renderer, Scene};
use std::env;
use std::process::exit;

fn main() {
    let args: Vec<String> = env::args().collect();
    if args.len() < 4 {
        eprintln!("usage: {} <scene.json> <integrator.json> <out.png>", args[0]);
        exit(2);
    }
    let run = || -> aqua_sys::Result<()> {
        let opts = LoadOptions { skip_missing: true, emission_from_mtl: true, ..Default::default() };
        let scene = Scene::load(&args[1], opts)?;
        let cfg = integrator_load(&args[2])?;
        let mut r = Renderer::new(&scene, &cfg, 0)?;
        let st = r.render(None, 0, cfg.spp)?;
        let (w, h) = r.size();
        let px = r.tonemap(0.0)?;
        write_png(&args[3], w, h, &px)?;
        println!(
            "{}x{} at {} spp: {:.3} ms, {} rays",
            w,
            h,
            cfg.spp,
            st.ms_frame,
            st.camera_rays + st.extension_rays + st.shadow_rays
        );
        Ok(())
    };
    if let Err(e) = run() {
        eprintln!("{}", e);
        exit(1);
    }
}
