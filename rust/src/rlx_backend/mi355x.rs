//! `mi355x` compute backend: `RlxBackend::knn_graph` on one AMD MI355X (gfx950)
//! through the C ABI of `libevoc_mi355x` (include/evoc_mi355x.h of the MI355X
//! kNN-graph library).  The other stages delegate to the strict reference, as
//! the RLX backends do (src/rlx_backend/delegate.rs).
//!
//! The library draws nothing from the caller's RNG itself: this backend makes
//! the reference's draws in the reference's order (src/knn.rs:416 and
//! :124-129) and passes them to `evoc_mi355x_knn_graph_seeded`, so the
//! caller's `NumpyRandomState` ends where the strict backend leaves it.
//! `strict_precision = true` is refused with the library's `NOT_BITWISE`
//! status (the trait's rule, src/rlx_backend/mod.rs:131-134): the GPU's f32
//! sums are not the strict backend's compiler-ordered fast-math sums, so the
//! caller selects `EVOC_BACKEND=strict` for bitwise results.

use super::{BackendError, ComputeBackend, RlxBackend};
use crate::knn::KnnError;
use crate::numpy_rng::NumpyRandomState;
use crate::{EmbeddingData, KnnGraphOptions};
use ndarray::Array2;
use sprs::CsMat;
use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};
use std::sync::Mutex;

// ---- include/evoc_mi355x.h, field for field (#[repr(C)], same order and widths) ----

#[repr(C)]
struct EvocKnnOptions {
    n_neighbors: i64,
    n_trees: i64,
    leaf_size: i64,
    max_candidates: i64,
    max_rptree_depth: i64,
    n_iters: i64,
    delta: f32,
    delta_improv: f32,
    use_sorted_updates: i32,
    deterministic: i32,
    brute_force_max_n: i64,
    mode: i32,
    precision: i32,
    margin: i32,
    strict_precision: i32,
    rayon_threads: i32,
    reserved0: i32,
}

#[repr(C)]
struct EvocKnnInput {
    dtype: i32,
    location: i32,
    data: *const c_void,
    n: i64,
    d: i64,
    ld: i64,
}

#[repr(C)]
struct EvocKnnOutput {
    location: i32,
    reserved0: i32,
    indices: *mut i32,
    distances: *mut f32,
    capacity_k: i64,
    k_out: i64,
    seconds: f64,
    device_seconds: f64,
    score_seconds: f64,
    iterations: i64,
    rescore_rounds: i64,
    offers_dropped: i64,
    queue_seconds: f64,
    score_pairs: i64,
    slabs: i64,
    join_bytes: i64,
    join_seconds: f64,
    join_launches: i64,
}

#[repr(C)]
struct EvocCtx {
    _opaque: [u8; 0],
}

const EVOC_DTYPE_F32: i32 = 0;
const EVOC_DTYPE_I8: i32 = 1;
const EVOC_DTYPE_U8: i32 = 2;
const EVOC_MEM_HOST: i32 = 0;
const EVOC_OK: c_int = 0;
const EVOC_ERR_EMPTY_DATA: c_int = 1;
const EVOC_ERR_INVALID_NEIGHBORS: c_int = 2;
const EVOC_ABI_VERSION: &str = "abi 6)";

#[link(name = "evoc_mi355x")]
extern "C" {
    fn evoc_mi355x_version() -> *const c_char;
    fn evoc_mi355x_create(device: i32, hbm_budget_bytes: u64, out: *mut *mut EvocCtx) -> c_int;
    fn evoc_mi355x_destroy(ctx: *mut EvocCtx);
    fn evoc_mi355x_last_error(ctx: *const EvocCtx) -> *const c_char;
    fn evoc_knn_options_default(opts: *mut EvocKnnOptions);
    fn evoc_knn_effective_k(n_neighbors: i64, n: i64) -> i64;
    fn evoc_knn_rng_draws(dtype: i32, n: i64, opts: *const EvocKnnOptions) -> i64;
    fn evoc_mi355x_knn_graph_seeded(
        ctx: *mut EvocCtx,
        input: *const EvocKnnInput,
        opts: *const EvocKnnOptions,
        tau_state: *const i64,
        tree_states: *const i64,
        n_tree_rows: i64,
        out: *mut EvocKnnOutput,
    ) -> c_int;
}

/// One device context (HIP stream, memory pool, HBM budget) per backend; calls
/// on it are serialised by the library and by this mutex.
pub struct Mi355xBackend {
    ctx: Mutex<*mut EvocCtx>,
}

// The context is used only under the mutex; the library is thread-safe per context.
unsafe impl Send for Mi355xBackend {}
unsafe impl Sync for Mi355xBackend {}

impl Mi355xBackend {
    /// A context on device `EVOC_MI355X_DEVICE` (default 0) with the library's
    /// default HBM budget (90 % of free device memory).
    pub fn new() -> Result<Self, BackendError> {
        let version = unsafe { CStr::from_ptr(evoc_mi355x_version()) }.to_string_lossy().into_owned();
        if !version.contains(EVOC_ABI_VERSION) {
            return Err(BackendError::Device {
                code: -1,
                message: format!("libevoc_mi355x `{version}` does not have the `{EVOC_ABI_VERSION}` this backend binds"),
            });
        }
        let device = std::env::var("EVOC_MI355X_DEVICE")
            .ok()
            .and_then(|v| v.parse::<i32>().ok())
            .unwrap_or(0);
        let mut ctx: *mut EvocCtx = std::ptr::null_mut();
        let rc = unsafe { evoc_mi355x_create(device, 0, &mut ctx) };
        if rc != EVOC_OK {
            return Err(BackendError::Device {
                code: rc,
                message: format!("evoc_mi355x_create(device {device}) failed (no gfx950 device visible?)"),
            });
        }
        Ok(Self { ctx: Mutex::new(ctx) })
    }
}

impl Drop for Mi355xBackend {
    fn drop(&mut self) {
        if let Ok(ctx) = self.ctx.lock() {
            unsafe { evoc_mi355x_destroy(*ctx) };
        }
    }
}

/// `KnnGraphOptions` (src/knn.rs:43-76) as the library's options; `None` fields
/// keep the library's derived defaults, which are the reference's (src/knn.rs:402-414).
fn c_options(opts: &KnnGraphOptions, strict_precision: bool) -> EvocKnnOptions {
    let mut o: EvocKnnOptions = unsafe { std::mem::zeroed() };
    unsafe { evoc_knn_options_default(&mut o) };
    o.n_neighbors = opts.n_neighbors as i64;
    o.n_trees = opts.n_trees.map_or(0, |t| t as i64);
    o.leaf_size = opts.leaf_size.map_or(0, |t| t as i64);
    o.max_candidates = opts.max_candidates.map_or(0, |t| t as i64);
    o.max_rptree_depth = opts.max_rptree_depth;
    o.n_iters = opts.n_iters.map_or(0, |t| t as i64);
    o.delta = opts.delta;
    o.delta_improv = opts.delta_improv;
    o.use_sorted_updates = opts.use_sorted_updates as i32;
    o.deterministic = opts.deterministic as i32;
    o.brute_force_max_n = opts.brute_force_max_n as i64;
    o.strict_precision = strict_precision as i32;
    // knn_graph runs in a 1-thread pool when deterministic (src/knn.rs:350-355),
    // and sizes its forest from the pool it runs in (src/knn.rs:404-407)
    o.rayon_threads = if opts.deterministic { 1 } else { rayon::current_num_threads() as i32 };
    o
}

impl RlxBackend for Mi355xBackend {
    fn kind(&self) -> ComputeBackend {
        ComputeBackend::Mi355x
    }

    fn knn_graph(
        &self,
        data: EmbeddingData,
        opts: KnnGraphOptions,
        rng: &mut NumpyRandomState,
        strict_precision: bool,
    ) -> Result<(Array2<i32>, Array2<f32>), KnnError> {
        // validation before any draw, as knn_graph_impl (src/knn.rs:382-393)
        if opts.n_neighbors == 0 {
            return Err(KnnError::InvalidNeighbors);
        }
        // standard-layout rows (a view the caller passed is copied once)
        let f32_rows;
        let i8_rows;
        let u8_rows;
        let (dtype, ptr, n, d) = match &data {
            EmbeddingData::Float32(a) => {
                f32_rows = a.as_standard_layout();
                (EVOC_DTYPE_F32, f32_rows.as_ptr() as *const c_void, a.nrows(), a.ncols())
            }
            EmbeddingData::Int8(a) => {
                i8_rows = a.as_standard_layout();
                (EVOC_DTYPE_I8, i8_rows.as_ptr() as *const c_void, a.nrows(), a.ncols())
            }
            EmbeddingData::UInt8(a) => {
                u8_rows = a.as_standard_layout();
                (EVOC_DTYPE_U8, u8_rows.as_ptr() as *const c_void, a.nrows(), a.ncols())
            }
        };
        if n == 0 {
            return Err(KnnError::EmptyData);
        }
        let o = c_options(&opts, strict_precision);
        // the reference's draws, in its order: randint3_for_tau (src/knn.rs:416), then
        // 3 per tree for make_forest (src/knn.rs:124-129), none on the brute-force branch
        let tau = rng.randint3_for_tau();
        let draws = unsafe { evoc_knn_rng_draws(dtype, n as i64, &o) };
        let trees: Vec<i64> = (0..(draws - 3).max(0))
            .map(|_| rng.randint(i32::MIN as i64 + 1, i32::MAX as i64 - 1))
            .collect();
        let k = unsafe { evoc_knn_effective_k(o.n_neighbors, n as i64) } as usize;
        let mut idx = Array2::<i32>::zeros((n, k));
        let mut dist = Array2::<f32>::zeros((n, k));
        let input = EvocKnnInput {
            dtype,
            location: EVOC_MEM_HOST,
            data: ptr,
            n: n as i64,
            d: d as i64,
            ld: d as i64,
        };
        let mut out: EvocKnnOutput = unsafe { std::mem::zeroed() };
        out.location = EVOC_MEM_HOST;
        out.indices = idx.as_mut_ptr();
        out.distances = dist.as_mut_ptr();
        out.capacity_k = k as i64;
        let ctx = self.ctx.lock().expect("mi355x context mutex poisoned");
        let rc = unsafe {
            evoc_mi355x_knn_graph_seeded(
                *ctx,
                &input,
                &o,
                tau.as_ptr(),
                if trees.is_empty() { std::ptr::null() } else { trees.as_ptr() },
                (trees.len() / 3) as i64,
                &mut out,
            )
        };
        match rc {
            EVOC_OK => Ok((idx, dist)),
            EVOC_ERR_EMPTY_DATA => Err(KnnError::EmptyData),
            EVOC_ERR_INVALID_NEIGHBORS => Err(KnnError::InvalidNeighbors),
            code => Err(KnnError::Backend(BackendError::Device {
                code,
                message: unsafe { CStr::from_ptr(evoc_mi355x_last_error(*ctx)) }
                    .to_string_lossy()
                    .into_owned(),
            })),
        }
    }

    fn label_propagation_init(
        &self,
        graph: &CsMat<f32>,
        n_label_prop_iter: usize,
        n_embedding_epochs: usize,
        approx_n_parts: usize,
        n_components: usize,
        scaling: f32,
        random_scale: f32,
        noise_level: f32,
        rng: &mut NumpyRandomState,
        data: Option<&Array2<f32>>,
        strict_precision: bool,
    ) -> Array2<f32> {
        super::strict::StrictBackend.label_propagation_init(
            graph,
            n_label_prop_iter,
            n_embedding_epochs,
            approx_n_parts,
            n_components,
            scaling,
            random_scale,
            noise_level,
            rng,
            data,
            strict_precision,
        )
    }

    fn node_embedding(
        &self,
        graph: &CsMat<f32>,
        n_components: usize,
        n_epochs: usize,
        initial_embedding: Option<Array2<f32>>,
        initial_alpha: f32,
        negative_sample_rate: f32,
        noise_level: f32,
        rng: &mut NumpyRandomState,
        reproducible_flag: bool,
        strict_precision: bool,
    ) -> Array2<f32> {
        super::strict::StrictBackend.node_embedding(
            graph,
            n_components,
            n_epochs,
            initial_embedding,
            initial_alpha,
            negative_sample_rate,
            noise_level,
            rng,
            reproducible_flag,
            strict_precision,
        )
    }
}
