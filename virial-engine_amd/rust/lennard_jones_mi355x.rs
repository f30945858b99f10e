//! MI355X (gfx950) back end of `virial_particles::lennard_jones`: the Rust
//! side of the drop-in, calling the C ABI of `libvirial_lj.so`
//! (`include/virial_lj.h`). It mirrors `LennardJones`
//! (crates/virial-particles/src/lennard_jones.rs:332-1375) call for call and
//! keeps that module's public types (`Particle`, `Settings`, `Thermostat`,
//! `Langevin`, `NoseHoover`, `Measurement`, `Profile`, `HeatExchange`,
//! `Drawn`); a maintainer adds it as
//! crates/virial-particles/src/lennard_jones_mi355x.rs next to them and links
//! the library from build.rs. The conversions `VljSettings::from(Settings)` and
//! `Measurement::from(&VljMeasurement)` are field-for-field copies (Thermostat
//! None / Langevin / NoseHoover as kind 0 / 1 / 2). This image has no Rust
//! toolchain, so the file is not compiled here; the library's own tests drive
//! the same ABI through ctypes.

// build.rs of virial-particles:
//     println!("cargo:rustc-link-search=native=<path to virial-engine_amd>");
//     println!("cargo:rustc-link-lib=dylib=virial_lj");

use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};

#[repr(C)] #[derive(Clone, Copy)] pub struct VljParticle { pub pos: [f64; 3], pub vel: [f64; 3] }
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct VljThermostat { pub kind: u32, pub temperature: f64, pub friction: f64, pub period: f64, pub seed: u64 }
#[repr(C)] #[derive(Clone, Copy)]
pub struct VljSettings {
    pub box_size: [f64; 3], pub cutoff: f64, pub dt: f64, pub skin: f64,
    pub thermostat: VljThermostat, pub shifted_force: i32, pub profile_bins: u32, pub precision: u32,
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct VljMeasurement {
    pub step: u64, pub time: f64, pub kinetic: f64, pub potential: f64, pub virial: f64,
    pub momentum: [f64; 3], pub count: u32, pub volume: f64, pub list_rebuilds: u32,
    pub list_overflow: u32, pub thermostat_energy: f64, pub condensed: u32,
    pub pairs_within_cutoff: u64, pub raw: [i64; 8],
}
#[repr(C)] #[derive(Clone, Copy, Default)]
pub struct VljTuneTrial { pub fill: f64, pub force_us: f64, pub build_us: f64, pub bricks: u32 }
#[repr(C)] pub struct VljSim { _private: [u8; 0] }

extern "C" {
    fn vlj_create(p: *const VljParticle, n: usize, s: *const VljSettings, out: *mut *mut VljSim) -> c_int;
    fn vlj_destroy(sim: *mut VljSim);
    fn vlj_step(sim: *mut VljSim, steps: u32) -> c_int;
    fn vlj_run(sim: *mut VljSim, steps: u64) -> c_int;
    fn vlj_measure(sim: *mut VljSim, accepted: *mut i32) -> c_int;
    fn vlj_poll(sim: *mut VljSim, out: *mut VljMeasurement, cap: usize, n: *mut usize) -> c_int;
    fn vlj_measure_now(sim: *mut VljSim, out: *mut VljMeasurement) -> c_int;
    fn vlj_read_particles(sim: *mut VljSim, out: *mut VljParticle, n: usize) -> c_int;
    fn vlj_set_temperature(sim: *mut VljSim, t: f64) -> c_int;
    fn vlj_box_size(sim: *const VljSim, out: *mut f64) -> c_int;
    fn vlj_sample_profile(sim: *mut VljSim) -> c_int;
    fn vlj_clear_profile(sim: *mut VljSim) -> c_int;
    fn vlj_read_profile(sim: *mut VljSim, out: *mut f64, cap: usize, samples: *mut u64) -> c_int;
    fn vlj_set_heat_exchange(sim: *mut VljSim, on: i32, hot: *const f64, cold: *const f64, energy: f64) -> c_int;
    fn vlj_heat_flux(sim: *const VljSim, flux: *mut f64, has: *mut i32) -> c_int;
    fn vlj_record_heat_current(sim: *mut VljSim, capacity: u32) -> c_int;
    fn vlj_sample_heat_current(sim: *mut VljSim) -> c_int;
    fn vlj_read_heat_current(sim: *mut VljSim, out: *mut f64, cap: usize, samples: *mut u32) -> c_int;
    fn vlj_last_error(sim: *const VljSim) -> *const c_char;
    fn vlj_render_buffers(sim: *mut VljSim, pos: *mut *mut c_void, vel: *mut *mut c_void) -> c_int;
    fn vlj_render(sim: *mut VljSim) -> c_int;
    fn vlj_set_drawn(sim: *mut VljSim, drawn: u32) -> c_int;
    fn vlj_save(sim: *mut VljSim, path: *const c_char) -> c_int;
    fn vlj_load(path: *const c_char, out: *mut *mut VljSim) -> c_int;
    // accessors (count(), volume(), steps(), time(), target_temperature())
    fn vlj_count(sim: *const VljSim) -> u32;
    fn vlj_volume(sim: *const VljSim) -> f64;
    fn vlj_steps(sim: *const VljSim) -> u64;
    fn vlj_time(sim: *const VljSim) -> f64;
    fn vlj_target_temperature(sim: *const VljSim, t: *mut f64, has: *mut i32) -> c_int;
    fn vlj_set_condensed_neighbours(sim: *mut VljSim, neighbours: u32) -> c_int;
    // the fixed-point sums as stored, for callers that combine them exactly
    fn vlj_read_profile_raw(sim: *mut VljSim, raw: *mut i64, cap: usize, samples: *mut u64) -> c_int;
    fn vlj_read_heat_current_raw(sim: *mut VljSim, out: *mut i64, cap: usize, samples: *mut u32) -> c_int;
    fn vlj_read_render(sim: *mut VljSim, pos: *mut f32, vel: *mut f32, n: usize) -> c_int;
    // the brick planner: measure three fills on the handle's own steps; the plan in use
    fn vlj_tune(sim: *mut VljSim, steps_per_candidate: u32, steps_taken: *mut u64) -> c_int;
    fn vlj_plan_info(sim: *const VljSim, fill: *mut f64, replans: *mut u32, trials: *mut VljTuneTrial,
                     capacity: usize, n_trials: *mut usize) -> c_int;
    // the batched form: a batch is a VljSim with several replicas
    fn vlj_batch_create(all: *const VljParticle, offsets: *const u64, s: *const VljSettings,
                        replicas: usize, out: *mut *mut VljSim) -> c_int;
    fn vlj_batch_run(b: *mut VljSim, steps: u64) -> c_int;
    fn vlj_batch_set_temperature(b: *mut VljSim, replica: usize, t: f64) -> c_int;
    fn vlj_batch_replicas(b: *const VljSim) -> usize;
    fn vlj_batch_measure_now(b: *mut VljSim, per_replica: *mut VljMeasurement, replicas: usize) -> c_int;
    fn vlj_batch_read_particles(b: *mut VljSim, out: *mut VljParticle, n: usize) -> c_int;
    fn vlj_batch_read_profile_raw(b: *mut VljSim, replica: usize, raw: *mut i64, cap: usize,
                                  samples: *mut u64) -> c_int;
    fn vlj_replica_info(sim: *const VljSim, replica: usize, s: *mut VljSettings, count: *mut u64,
                        replicas: *mut usize) -> c_int;
    // initial states (lattice(), fcc(), random insertion) and the fixed-point sum rule
    fn vlj_lattice(n: usize, box_size: f64, t: f64, seed: u64, out: *mut VljParticle) -> c_int;
    fn vlj_lattice_box(n: usize, size: *const f64, t: f64, seed: u64, out: *mut VljParticle) -> c_int;
    fn vlj_fcc(cells: u32, density: f64, t: f64, seed: u64, out: *mut VljParticle, side: *mut f64) -> c_int;
    fn vlj_fcc_count(n: usize, density: f64, t: f64, seed: u64, out: *mut VljParticle, side: *mut f64) -> c_int;
    fn vlj_random_insertion(n: usize, bx: *const f64, min_distance: f64, t: f64, seed: u64,
                            out: *mut VljParticle) -> c_int;
    fn vlj_fixed_sum(values: *const f64, n: usize, fixed_scale: f64, out: *mut i64) -> c_int;
    // the same sum on device pointers, allocating nothing (for callers that cannot allocate)
    fn vlj_fixed_sum_device(values: *const f64, n: usize, fixed_scale: f64, out: *mut i64,
                            stream: *mut c_void) -> c_int;
}
// Not bound: vlj_debug_state, vlj_debug_cells, vlj_debug_lists, vlj_debug_grid,
// vlj_debug_counters, vlj_debug_replica_counters, vlj_debug_params, vlj_debug_plan and
// vlj_debug_parts (planning overrides),
// vlj_debug_swizzle (the XCD table mapping), vlj_debug_rcp and vlj_debug_rcp32
// (the force reciprocals), vlj_copy_bandwidth, vlj_fp64_fma_rate and
// vlj_timed_run are test and benchmark hooks with no counterpart in the
// reference; the library's own tests and bench.py call them.

/// Same contract as the wgpu LennardJones: errors panic with the reference's messages.
pub struct LennardJones { sim: *mut VljSim, count: usize }

impl LennardJones {
    fn check(&self, code: c_int) {
        if code != 0 {
            let msg = unsafe { CStr::from_ptr(vlj_last_error(self.sim)) };
            panic!("{}", msg.to_string_lossy());
        }
    }

    /// lennard_jones.rs:443; the wgpu device and queue are replaced by the current HIP device.
    pub fn new(particles: &[Particle], settings: Settings) -> Self {
        let ps: Vec<VljParticle> = particles.iter().map(|p| VljParticle { pos: p.pos, vel: p.vel }).collect();
        let s = VljSettings::from(settings); // maps Thermostat::{None, Langevin, NoseHoover} to kind 0/1/2
        let mut sim = std::ptr::null_mut();
        let code = unsafe { vlj_create(ps.as_ptr(), ps.len(), &s, &mut sim) };
        if code != 0 {
            let msg = unsafe { CStr::from_ptr(vlj_last_error(std::ptr::null())) };
            panic!("{}", msg.to_string_lossy());
        }
        Self { sim, count: ps.len() }
    }

    /// lennard_jones.rs:1092: the encoder argument goes away; the steps are
    /// enqueued on the library's stream.
    pub fn step(&mut self, steps: u32) { self.check(unsafe { vlj_step(self.sim, steps) }) }
    pub fn run(&mut self, steps: u64) { self.check(unsafe { vlj_run(self.sim, steps) }) }
    pub fn measure(&mut self) -> bool {
        let mut ok = 0;
        self.check(unsafe { vlj_measure(self.sim, &mut ok) });
        ok != 0
    }
    pub fn after_submit(&mut self) {}
    pub fn poll(&mut self) -> Vec<Measurement> {
        let mut out = [VljMeasurement::default(); 3];
        let mut n = 0;
        self.check(unsafe { vlj_poll(self.sim, out.as_mut_ptr(), 3, &mut n) });
        out[..n].iter().map(Measurement::from).collect()
    }
    pub fn measure_now(&mut self) -> Measurement {
        let mut m = VljMeasurement::default();
        self.check(unsafe { vlj_measure_now(self.sim, &mut m) });
        Measurement::from(&m)
    }
    pub fn read_particles(&self) -> Vec<Particle> {
        let mut out = vec![VljParticle { pos: [0.0; 3], vel: [0.0; 3] }; self.count];
        self.check(unsafe { vlj_read_particles(self.sim, out.as_mut_ptr(), out.len()) });
        out.iter().map(|p| Particle { pos: p.pos, vel: p.vel }).collect()
    }
    pub fn set_temperature(&mut self, t: f64) { self.check(unsafe { vlj_set_temperature(self.sim, t) }) }

    /// lennard_jones.rs:1167-1325: profiles along z. The reference's Profile
    /// is filled from the averaged sums (8 per slab, then the whole-box W_aa).
    pub fn sample_profile(&mut self) { self.check(unsafe { vlj_sample_profile(self.sim) }) }
    pub fn clear_profile(&mut self) { self.check(unsafe { vlj_clear_profile(self.sim) }) }
    pub fn read_profile(&self, bins: usize, box_size: [f64; 3]) -> Profile {
        let mut sums = vec![0.0; bins * 8 + 3];
        let mut samples = 0u64;
        self.check(unsafe { vlj_read_profile(self.sim, sums.as_mut_ptr(), sums.len(), &mut samples) });
        Profile::from_sums(&sums, samples, box_size) // slab b: sums[8b..8b+8]
    }

    /// lennard_jones.rs:927-1012: heat exchange and the heat-current series.
    pub fn set_heat_exchange(&mut self, h: Option<HeatExchange>) {
        let (on, hot, cold, e) = match h {
            Some(h) => (1, [h.hot.0, h.hot.1], [h.cold.0, h.cold.1], h.energy),
            None => (0, [0.0; 2], [0.0; 2], 0.0),
        };
        self.check(unsafe { vlj_set_heat_exchange(self.sim, on, hot.as_ptr(), cold.as_ptr(), e) })
    }
    pub fn heat_flux(&self) -> Option<f64> {
        let (mut f, mut has) = (0.0, 0);
        self.check(unsafe { vlj_heat_flux(self.sim, &mut f, &mut has) });
        (has != 0).then_some(f)
    }
    pub fn record_heat_current(&mut self, capacity: u32) {
        self.check(unsafe { vlj_record_heat_current(self.sim, capacity) })
    }
    pub fn sample_heat_current(&mut self) { self.check(unsafe { vlj_sample_heat_current(self.sim) }) }
    pub fn read_heat_current(&self, capacity: usize) -> Vec<[f64; 3]> {
        let mut out = vec![0.0; 3 * capacity];
        let mut n = 0u32;
        self.check(unsafe { vlj_read_heat_current(self.sim, out.as_mut_ptr(), capacity, &mut n) });
        out.chunks(3).take(n as usize).map(|c| [c[0], c[1], c[2]]).collect()
    }

    /// lennard_jones.rs:1041-1051, 885-889: the render arrays, float4 per
    /// particle on the device (slot order). A renderer on the same GPU imports
    /// the two allocations (HIP-Vulkan interop, or a copy into its own buffers
    /// after `render`); each measurement refreshes them as the reference's
    /// measurement refreshes the w channel.
    pub fn render_buffers(&mut self) -> (*mut c_void, *mut c_void) {
        let (mut p, mut v) = (std::ptr::null_mut(), std::ptr::null_mut());
        self.check(unsafe { vlj_render_buffers(self.sim, &mut p, &mut v) });
        (p, v)
    }
    pub fn render(&mut self) { self.check(unsafe { vlj_render(self.sim) }) }
    pub fn set_drawn(&mut self, drawn: Drawn) { self.check(unsafe { vlj_set_drawn(self.sim, drawn as u32) }) }

    /// Checkpoint and exact resume (no counterpart in the reference).
    pub fn save(&mut self, path: &std::path::Path) {
        let c = std::ffi::CString::new(path.to_string_lossy().as_bytes()).unwrap();
        self.check(unsafe { vlj_save(self.sim, c.as_ptr()) })
    }
    pub fn load(path: &std::path::Path) -> Self {
        let c = std::ffi::CString::new(path.to_string_lossy().as_bytes()).unwrap();
        let mut sim = std::ptr::null_mut();
        if unsafe { vlj_load(c.as_ptr(), &mut sim) } != 0 {
            let msg = unsafe { CStr::from_ptr(vlj_last_error(std::ptr::null())) };
            panic!("{}", msg.to_string_lossy());
        }
        let count = unsafe { vlj_count(sim) } as usize;
        Self { sim, count }
    }

    /// The brick planner (no counterpart in the reference: the wgpu kernels
    /// have no tiles to plan). Runs `steps_per_candidate` steps under each
    /// candidate fill and keeps the fastest; returns the steps it ran.
    pub fn tune(&mut self, steps_per_candidate: u32) -> u64 {
        let mut taken = 0;
        self.check(unsafe { vlj_tune(self.sim, steps_per_candidate, &mut taken) });
        taken
    }
    pub fn plan_info(&self) -> (f64, u32, Vec<VljTuneTrial>) {
        let (mut fill, mut replans, mut n) = (0.0, 0u32, 0usize);
        let mut trials = [VljTuneTrial::default(); 8];
        self.check(unsafe { vlj_plan_info(self.sim, &mut fill, &mut replans, trials.as_mut_ptr(), 8, &mut n) });
        (fill, replans, trials[..n.min(8)].to_vec())
    }
}

impl Drop for LennardJones {
    fn drop(&mut self) { unsafe { vlj_destroy(self.sim) } }
}

/// Independent replicas stepped together: one launch of each kernel over the
/// replica-offset table. Replica r's particles are `particles[offsets[r]..offsets[r + 1]]`.
pub struct Batch { sim: *mut VljSim, replicas: usize, count: usize }

impl Batch {
    fn check(&self, code: c_int) {
        if code != 0 {
            let msg = unsafe { CStr::from_ptr(vlj_last_error(self.sim)) };
            panic!("{}", msg.to_string_lossy());
        }
    }
    pub fn new(replicas: &[(&[Particle], Settings)]) -> Self {
        let mut all = Vec::new();
        let mut offsets = vec![0u64];
        let mut settings = Vec::new();
        for (ps, s) in replicas {
            all.extend(ps.iter().map(|p| VljParticle { pos: p.pos, vel: p.vel }));
            offsets.push(all.len() as u64);
            settings.push(VljSettings::from(s.clone()));
        }
        let mut sim = std::ptr::null_mut();
        let code = unsafe { vlj_batch_create(all.as_ptr(), offsets.as_ptr(), settings.as_ptr(), replicas.len(), &mut sim) };
        if code != 0 {
            let msg = unsafe { CStr::from_ptr(vlj_last_error(std::ptr::null())) };
            panic!("{}", msg.to_string_lossy());
        }
        Self { sim, replicas: replicas.len(), count: all.len() }
    }
    pub fn run(&mut self, steps: u64) { self.check(unsafe { vlj_batch_run(self.sim, steps) }) }
    pub fn set_temperature(&mut self, replica: usize, t: f64) {
        self.check(unsafe { vlj_batch_set_temperature(self.sim, replica, t) })
    }
    pub fn measure_now(&mut self) -> Vec<Measurement> {
        let mut out = vec![VljMeasurement::default(); self.replicas];
        self.check(unsafe { vlj_batch_measure_now(self.sim, out.as_mut_ptr(), out.len()) });
        out.iter().map(Measurement::from).collect()
    }
    pub fn read_particles(&self) -> Vec<Particle> {
        let mut out = vec![VljParticle { pos: [0.0; 3], vel: [0.0; 3] }; self.count];
        self.check(unsafe { vlj_batch_read_particles(self.sim, out.as_mut_ptr(), out.len()) });
        out.iter().map(|p| Particle { pos: p.pos, vel: p.vel }).collect()
    }
    pub fn tune(&mut self, steps_per_candidate: u32) -> u64 {
        let mut taken = 0;
        self.check(unsafe { vlj_tune(self.sim, steps_per_candidate, &mut taken) });
        taken
    }
}

impl Drop for Batch {
    fn drop(&mut self) { unsafe { vlj_destroy(self.sim) } }
}
