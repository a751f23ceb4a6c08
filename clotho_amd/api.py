"""ctypes binding of the C ABI declared in include/clotho_amd.h.

The Python layer is a thin host-side mirror for tests and benchmarks; all
compute runs in libclotho_amd.so (HIP kernels for gfx950).  Loading fails
loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libclotho_amd.so")

FIT_CONSTANT = 0
FIT_GAUSSIAN = 1
FIT_QUADRATIC = 2
FIT_GAUSSIAN_FIXED = 3
FIT_MV_STABILIZING = 4

TRAITS_PLAIN_FP32 = 0
TRAITS_MFMA_BF16 = 1
TRAITS_EXACT_FP64 = 2
TRAITS_WALK_FP32 = 3
TRAITS_EXACT_I8 = 4

FLAG_HOST_MUTATIONS = 1

ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "capacity exceeded", -4: "bad state"}


class Config(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("replicate", C.c_uint32),
        ("n_individuals", C.c_uint32),
        ("trait_count", C.c_uint32),
        ("fitness_kind", C.c_uint32),
        ("mu", C.c_double),
        ("rho", C.c_double),
        ("bias", C.c_double),
        ("neutral_p", C.c_double),
        ("trait_mean", C.c_double),
        ("trait_sigma", C.c_double),
        ("fitness_scale", C.c_double),
        ("env_sigma", C.c_double),
        ("effect_quantum_log2", C.c_int32),
        ("max_alleles", C.c_uint32),
        ("compact_every", C.c_uint32),
        ("max_individuals", C.c_uint32),
        ("intermediates", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class Shape(C.Structure):
    _fields_ = [
        ("generation", C.c_uint32),
        ("n_individuals", C.c_uint32),
        ("haplotypes", C.c_uint32),
        ("words", C.c_uint32),
        ("alleles", C.c_uint32),
        ("new_mutations", C.c_uint32),
        ("fixed_count", C.c_uint32),
        ("lost_count", C.c_uint32),
        ("free_count", C.c_uint32),
        ("variable_count", C.c_uint32),
        ("all_neutral", C.c_uint32),
        ("fitness_constant", C.c_uint32),
        ("error", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def make_config(n, mu, rho, seed=42, traits=1, fitness=FIT_CONSTANT, bias=0.5, neutral_p=0.5,
                mean=0.0, sigma=1.0, fitness_scale=1.0, env_sigma=0.0, quantum_log2=-24,
                max_alleles=1 << 20, replicate=0, compact=False, max_individuals=0,
                intermediates=0, host_mutations=False):
    """Build a Config; defaults follow the reference's parameter defaults
    (sequence_bias.p 0.5, neutral.p 0.5, traits.mean 0, traits.sigma 1)."""
    c = Config()
    c.seed = seed
    c.replicate = replicate
    c.n_individuals = n
    c.trait_count = traits
    c.fitness_kind = fitness
    c.mu = mu
    c.rho = rho
    c.bias = bias
    c.neutral_p = neutral_p
    c.trait_mean = mean
    c.trait_sigma = sigma
    c.fitness_scale = fitness_scale
    c.env_sigma = env_sigma
    c.effect_quantum_log2 = quantum_log2
    c.max_alleles = max_alleles
    c.compact_every = 1 if compact else 0
    c.max_individuals = max_individuals
    c.intermediates = intermediates
    c.flags = FLAG_HOST_MUTATIONS if host_mutations else 0
    return c


_lib = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

EXPORTS = [
    "clotho_create", "clotho_destroy", "clotho_last_error", "clotho_version",
    "clotho_poisson_table", "clotho_set_population", "clotho_get_shape",
    "clotho_download_population", "clotho_download_alleles", "clotho_download_free_space",
    "clotho_download_traits", "clotho_download_fitness", "clotho_download_gametes",
    "clotho_run", "clotho_generation", "clotho_phase_draw", "clotho_phase_recombine_mutate",
    "clotho_phase_traits",
    "clotho_phase_free_space", "clotho_phase_fitness", "clotho_phase_commit", "clotho_sync",
    "clotho_timings", "clotho_set_timing", "clotho_trait_contract",
    "clotho_set_synthetic_population", "clotho_stream_copy", "clotho_trait_contract_synthetic",
    "clotho_autotune",
    "clotho_batch_create", "clotho_batch_destroy", "clotho_batch_last_error",
    "clotho_batch_size", "clotho_batch_replicate", "clotho_batch_run", "clotho_batch_timings",
    "clotho_batch_set_timing", "clotho_compact", "clotho_device_bytes",
    "clotho_sample_rows", "clotho_allele_counts", "clotho_set_population_size",
    "clotho_pairwise_differences", "clotho_device_memory",
    "clotho_set_parents", "clotho_add_mutations", "clotho_evaluate_fitness",
    "clotho_crossover_rows", "clotho_classify_free_space",
    "clotho_trait_contract_exact_noalloc", "clotho_get_launch_shape", "clotho_batch_autotune",
    "clotho_set_launch_shape",
]

BATCH_MODES = {"graph": 0, "eager": 1, "sequential": 2}


def lib():
    """Load libclotho_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.clotho_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.clotho_create.restype = C.c_int
    L.clotho_destroy.argtypes = [vp]
    L.clotho_last_error.argtypes = [vp]
    L.clotho_last_error.restype = C.c_char_p
    L.clotho_version.restype = C.c_char_p
    L.clotho_poisson_table.argtypes = [C.c_double, f64p, C.c_int]
    L.clotho_poisson_table.restype = C.c_int
    L.clotho_set_population.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp]
    L.clotho_set_synthetic_population.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_uint64]
    L.clotho_stream_copy.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double), C.c_int]
    L.clotho_get_shape.argtypes = [vp, vp]
    L.clotho_download_population.argtypes = [vp, vp, C.c_size_t]
    L.clotho_download_alleles.argtypes = [vp, vp, vp, vp, C.c_size_t]
    L.clotho_download_free_space.argtypes = [vp, vp, vp, C.c_size_t]
    L.clotho_download_traits.argtypes = [vp, vp, C.c_size_t]
    L.clotho_download_fitness.argtypes = [vp, vp, C.c_size_t]
    L.clotho_download_gametes.argtypes = [vp, vp, vp, vp, vp, C.c_size_t]
    L.clotho_run.argtypes = [vp, C.c_uint32, C.c_int]
    L.clotho_set_population_size.argtypes = [vp, C.c_uint32]
    L.clotho_pairwise_differences.argtypes = [vp, u32p, C.c_uint32, u32p, C.c_size_t,
                                              C.POINTER(C.c_double)]
    for n in ("clotho_generation", "clotho_phase_draw", "clotho_phase_recombine_mutate",
              "clotho_phase_traits",
              "clotho_phase_free_space", "clotho_phase_fitness", "clotho_phase_commit",
              "clotho_sync", "clotho_compact"):
        getattr(L, n).argtypes = [vp]
    L.clotho_timings.argtypes = [vp, f64p, C.c_int]
    L.clotho_autotune.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.clotho_get_launch_shape.argtypes = [vp, vp]
    L.clotho_set_launch_shape.argtypes = [vp, vp]
    L.clotho_batch_autotune.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.clotho_set_timing.argtypes = [vp, C.c_int]
    L.clotho_set_timing.restype = None
    L.clotho_trait_contract.argtypes = [vp, C.c_uint32, C.c_uint32, vp, C.c_uint32, C.c_uint32,
                                        C.c_int, C.c_int, C.c_int, vp, C.POINTER(C.c_double),
                                        C.c_int]
    L.clotho_trait_contract_synthetic.argtypes = [C.c_uint32, C.c_uint32, vp, C.c_uint64, vp,
                                                  C.c_uint32, C.c_int, C.c_int, C.c_int, vp,
                                                  C.POINTER(C.c_double), C.c_int]
    L.clotho_device_bytes.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.clotho_device_memory.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.clotho_classify_free_space.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp,
                                             C.c_int]
    L.clotho_trait_contract_exact_noalloc.argtypes = [vp, C.c_uint32, C.c_uint32, vp,
                                                      C.c_uint32, C.c_uint32, vp, vp]
    L.clotho_set_parents.argtypes = [vp, vp, C.c_uint32]
    L.clotho_add_mutations.argtypes = [vp, vp, vp, vp, C.c_uint32, vp]
    L.clotho_evaluate_fitness.argtypes = [vp]
    L.clotho_crossover_rows.argtypes = [vp, vp, vp, C.c_uint32, vp, vp, C.c_uint32, C.c_uint32,
                                        vp, C.c_uint32]
    L.clotho_sample_rows.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     u32p]
    L.clotho_allele_counts.argtypes = [vp, vp, C.c_uint32, u32p, C.c_size_t, u32p,
                                       C.POINTER(C.c_double)]
    L.clotho_batch_create.argtypes = [vp, C.c_uint32, C.c_int, C.POINTER(vp)]
    L.clotho_batch_destroy.argtypes = [vp]
    L.clotho_batch_destroy.restype = None
    L.clotho_batch_last_error.argtypes = [vp]
    L.clotho_batch_last_error.restype = C.c_char_p
    L.clotho_batch_size.argtypes = [vp]
    L.clotho_batch_size.restype = C.c_uint32
    L.clotho_batch_replicate.argtypes = [vp, C.c_uint32]
    L.clotho_batch_replicate.restype = vp
    L.clotho_batch_run.argtypes = [vp, C.c_uint32, C.c_int, C.POINTER(C.c_double)]
    L.clotho_batch_timings.argtypes = [vp, f64p, C.c_int]
    L.clotho_batch_set_timing.argtypes = [vp, C.c_int]
    L.clotho_batch_set_timing.restype = None
    _lib = L
    return L


class ClothoError(RuntimeError):
    pass


class Simulation:
    """One context: a population pair and its allele table on one GPU."""

    def __init__(self, cfg: Config, device: int = 0):
        self.cfg = cfg
        self.ctx = C.c_void_p()
        rc = lib().clotho_create(C.addressof(cfg), device, C.byref(self.ctx))
        if rc != 0:
            raise ClothoError(f"clotho_create failed ({ERRORS.get(rc, rc)})")

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().clotho_last_error(self.ctx)
            raise ClothoError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self):
        if self.ctx:
            lib().clotho_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_population(self, rows, positions, effects, fitness=None):
        rows = np.ascontiguousarray(rows, np.uint64)
        positions = np.ascontiguousarray(positions, np.float64)
        effects = np.ascontiguousarray(effects, np.float64)
        fp = None
        if fitness is not None:
            fitness = np.ascontiguousarray(fitness, np.float64)
            fp = fitness.ctypes.data
        self._check(lib().clotho_set_population(self.ctx, rows.ctypes.data, rows.shape[1],
                                                positions.ctypes.data, effects.ctypes.data,
                                                positions.size, fp), "set_population")

    def set_synthetic(self, freqs, positions, effects, seed=42):
        """Standing variation drawn on the device: row r has slot j with
        probability freqs[j] (see clotho_set_synthetic_population)."""
        freqs = np.ascontiguousarray(freqs, np.float64)
        positions = np.ascontiguousarray(positions, np.float64)
        effects = np.ascontiguousarray(effects, np.float64)
        self._check(lib().clotho_set_synthetic_population(
            self.ctx, freqs.size, freqs.ctypes.data, positions.ctypes.data,
            effects.ctypes.data, seed), "set_synthetic_population")

    def run(self, n=1, graph=False):
        self._check(lib().clotho_run(self.ctx, n, 1 if graph else 0), "run")

    def generation(self):
        self._check(lib().clotho_generation(self.ctx), "generation")

    def phase(self, name):
        self._check(getattr(lib(), "clotho_phase_" + name)(self.ctx), name)

    def sync(self):
        self._check(lib().clotho_sync(self.ctx), "sync")

    def shape(self) -> Shape:
        s = Shape()
        self._check(lib().clotho_get_shape(self.ctx, C.addressof(s)), "get_shape")
        return s

    def rows(self):
        s = self.shape()
        out = np.zeros((s.haplotypes, s.words), np.uint64)
        self._check(lib().clotho_download_population(self.ctx, out.ctypes.data, out.size),
                    "download_population")
        return out

    def alleles(self):
        s = self.shape()
        t = self.cfg.trait_count
        pos = np.zeros(max(s.alleles, 1), np.float64)
        eff = np.zeros(max(s.alleles * t, 1), np.float64)
        age = np.zeros(max(s.alleles, 1), np.uint32)
        self._check(lib().clotho_download_alleles(self.ctx, pos.ctypes.data, eff.ctypes.data,
                                                  age.ctypes.data, s.alleles),
                    "download_alleles")
        return pos[: s.alleles], eff[: s.alleles * t].reshape(s.alleles, t), age[: s.alleles]

    def free_space(self):
        s = self.shape()
        fx = np.zeros(max(s.fixed_count, 1), np.uint32)
        ls = np.zeros(max(s.lost_count, 1), np.uint32)
        self._check(lib().clotho_download_free_space(self.ctx, fx.ctypes.data, ls.ctypes.data,
                                                     max(s.fixed_count, s.lost_count, 1)),
                    "download_free_space")
        return fx[: s.fixed_count], ls[: s.lost_count]

    def traits(self):
        s = self.shape()
        out = np.zeros((s.haplotypes, self.cfg.trait_count), np.float64)
        self._check(lib().clotho_download_traits(self.ctx, out.ctypes.data, out.size),
                    "download_traits")
        return out

    def fitness(self):
        s = self.shape()
        out = np.zeros(s.n_individuals, np.float64)
        self._check(lib().clotho_download_fitness(self.ctx, out.ctypes.data, out.size),
                    "download_fitness")
        return out

    def gametes(self):
        s = self.shape()
        h = s.haplotypes
        top, bot, ev, mu = (np.zeros(h, np.uint32) for _ in range(4))
        self._check(lib().clotho_download_gametes(self.ctx, top.ctypes.data, bot.ctypes.data,
                                                  ev.ctypes.data, mu.ctypes.data, h),
                    "download_gametes")
        return top, bot, ev, mu

    def device_bytes(self):
        out = C.c_uint64(0)
        self._check(lib().clotho_device_bytes(self.ctx, C.byref(out)), "device_bytes")
        return out.value

    def allele_counts(self, rows=None, timed=False):
        """(column_counts[alleles], row_counts) of the current population over
        `rows` (all haplotypes when None)."""
        s = self.shape()
        n = s.haplotypes if rows is None else len(rows)
        cols = np.zeros(max(s.alleles, 1), np.uint32)
        rc = np.zeros(max(n, 1), np.uint32)
        ms = C.c_double(0.0)
        rp = None
        if rows is not None:
            rows = np.ascontiguousarray(rows, np.uint32)
            rp = rows.ctypes.data
        self._check(lib().clotho_allele_counts(self.ctx, rp, n, cols, cols.size, rc,
                                               C.byref(ms)), "allele_counts")
        out = (cols[: s.alleles], rc[:n])
        return (out + (ms.value,)) if timed else out

    def compact(self):
        self._check(lib().clotho_compact(self.ctx), "compact")

    def pairwise_differences(self, rows, timed=False):
        """popcount(row a xor row b) over the current rows for the pairs a < b
        of `rows` (in the given order), packed row by row."""
        rows = np.ascontiguousarray(rows, np.uint32)
        n = rows.size
        out = np.zeros(max(1, n * (n - 1) // 2), np.uint32)
        ms = C.c_double()
        self._check(lib().clotho_pairwise_differences(self.ctx, rows, n, out, out.size,
                                                      C.byref(ms)), "pairwise_differences")
        out = out[: n * (n - 1) // 2]
        return (out, ms.value) if timed else out

    def set_population_size(self, n):
        """Offspring individuals of every following generation (the pN of a
        population-growth model); at most max(N, max_individuals)."""
        self._check(lib().clotho_set_population_size(self.ctx, int(n)), "set_population_size")

    def autotune(self):
        """Pick the recombination row-chunk count by timing (state unchanged)."""
        ch = C.c_uint32(0)
        self._check(lib().clotho_autotune(self.ctx, C.byref(ch)), "autotune")
        return ch.value

    def set_parents(self, parents):
        """Parent individual of every gamete of the next generation."""
        parents = np.ascontiguousarray(parents, np.uint32)
        self._check(lib().clotho_set_parents(self.ctx, parents.ctypes.data, parents.size),
                    "set_parents")

    def add_mutations(self, rows, positions, effects):
        """New alleles on the current population; returns their slots."""
        rows = np.ascontiguousarray(rows, np.uint32)
        positions = np.ascontiguousarray(positions, np.float64)
        effects = np.ascontiguousarray(effects, np.float64)
        slots = np.zeros(max(rows.size, 1), np.uint32)
        self._check(lib().clotho_add_mutations(self.ctx, rows.ctypes.data, positions.ctypes.data,
                                               effects.ctypes.data, rows.size, slots.ctypes.data),
                    "add_mutations")
        return slots[: rows.size]

    def evaluate_fitness(self):
        self._check(lib().clotho_evaluate_fitness(self.ctx), "evaluate_fitness")

    def crossover_rows(self, top, bottom, n_events, event_keys, out_words=None):
        """Per-gamete recombination of rows top[g] / bottom[g] with event keys
        event_keys[g, :n_events[g]] against the current allele positions."""
        top = np.ascontiguousarray(top, np.uint64)
        bottom = np.ascontiguousarray(bottom, np.uint64)
        n_events = np.ascontiguousarray(n_events, np.uint32)
        event_keys = np.ascontiguousarray(event_keys, np.uint32)
        g, w = top.shape
        ow = w if out_words is None else out_words
        out = np.zeros((g, ow), np.uint64)
        stride = event_keys.shape[1] if event_keys.ndim == 2 else 0
        self._check(lib().clotho_crossover_rows(self.ctx, top.ctypes.data, bottom.ctypes.data, w,
                                                n_events.ctypes.data, event_keys.ctypes.data,
                                                stride, g, out.ctypes.data, ow), "crossover_rows")
        return out

    def launch_shape(self):
        """The launch shapes in use (dict of clotho_launch_shape)."""
        out = (C.c_uint32 * 8)()
        self._check(lib().clotho_get_launch_shape(self.ctx, out), "get_launch_shape")
        keys = ("recombine_chunks", "traits_variant", "traits_variants", "compact_rows",
                "compact_chunks", "compact_tile_words")
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def set_launch_shape(self, **shape):
        """Apply launch shapes (keys of launch_shape(); missing keys keep theirs)."""
        cur = self.launch_shape()
        cur.update(shape)
        arr = (C.c_uint32 * 8)(cur["recombine_chunks"], cur["traits_variant"],
                               cur["traits_variants"], cur["compact_rows"],
                               cur["compact_chunks"], cur["compact_tile_words"])
        self._check(lib().clotho_set_launch_shape(self.ctx, arr), "set_launch_shape")

    def set_timing(self, on=True):
        lib().clotho_set_timing(self.ctx, 1 if on else 0)

    def timings(self):
        out = np.zeros(8, np.float64)
        n = lib().clotho_timings(self.ctx, out, out.size)
        return n, out


def sample_rows(seed, generation, replicate, population, k):
    out = np.zeros(max(min(k, population), 1), np.uint32)
    rc = lib().clotho_sample_rows(seed, generation, replicate, population, k, out)
    if rc != 0:
        raise ClothoError(f"sample_rows failed ({ERRORS.get(rc, rc)})")
    return out[: min(k, population)]


def poisson_table(lam, cap=128):
    out = np.zeros(cap, np.float64)
    n = lib().clotho_poisson_table(float(lam), out, cap)
    if n < 0:
        raise ClothoError("poisson table: lambda out of range")
    return out[:n].copy()


def stream_copy_gbps(nbytes=1 << 30, repeats=20, device=0):
    """Measured device copy bandwidth (read + write bytes per second / 1e9)."""
    out = C.c_double(0.0)
    rc = lib().clotho_stream_copy(nbytes, repeats, C.byref(out), device)
    if rc != 0:
        raise ClothoError(f"stream_copy failed ({ERRORS.get(rc, rc)})")
    return out.value


def classify_free_space(fixed_and, variable_or, max_alleles, device=0):
    """(fixed, lost) slots of reduced AND / OR column masks, below max_alleles."""
    fa = np.ascontiguousarray(fixed_and, np.uint64)
    vo = np.ascontiguousarray(variable_or, np.uint64)
    cap = max(1, min(fa.size * 64, max_alleles))
    fixed = np.zeros(cap, np.uint32)
    lost = np.zeros(cap, np.uint32)
    counts = np.zeros(2, np.uint32)
    rc = lib().clotho_classify_free_space(fa.ctypes.data, vo.ctypes.data, fa.size, max_alleles,
                                          fixed.ctypes.data, lost.ctypes.data,
                                          counts.ctypes.data, device)
    if rc != 0:
        raise ClothoError(f"classify_free_space failed ({ERRORS.get(rc, rc)})")
    return fixed[: counts[0]], lost[: counts[1]]


def device_memory(device=0):
    """(free, total) device memory in bytes."""
    f, t = C.c_uint64(0), C.c_uint64(0)
    rc = lib().clotho_device_memory(device, C.byref(f), C.byref(t))
    if rc != 0:
        raise ClothoError(f"device_memory failed ({ERRORS.get(rc, rc)})")
    return f.value, t.value


def trait_contract(rows, effects, mode, split_terms=3, repeats=1, device=0):
    """z = genotypes x effects on the GPU; returns (z [H,T] float64, avg kernel ms)."""
    rows = np.ascontiguousarray(rows, np.uint64)
    effects = np.ascontiguousarray(effects, np.float64)
    h, w = rows.shape
    s, t = effects.shape
    z = np.zeros((h, t), np.float64)
    ms = C.c_double(0.0)
    rc = lib().clotho_trait_contract(rows.ctypes.data, h, w, effects.ctypes.data, s, t, mode,
                                     split_terms, repeats, z.ctypes.data, C.byref(ms), device)
    if rc != 0:
        raise ClothoError(f"trait_contract failed ({ERRORS.get(rc, rc)})")
    return z, ms.value


def trait_contract_synthetic(h, freqs, effects, mode, seed=42, split_terms=3, repeats=1,
                             device=0):
    """Contraction over genotypes drawn on the device (see
    clotho_trait_contract_synthetic); returns (z [H,T] float64, avg kernel ms)."""
    freqs = np.ascontiguousarray(freqs, np.float64)
    effects = np.ascontiguousarray(effects, np.float64)
    s, t = effects.shape
    assert freqs.size == s
    z = np.zeros((h, t), np.float64)
    ms = C.c_double(0.0)
    rc = lib().clotho_trait_contract_synthetic(h, s, freqs.ctypes.data, seed,
                                               effects.ctypes.data, t, mode, split_terms,
                                               repeats, z.ctypes.data, C.byref(ms), device)
    if rc != 0:
        raise ClothoError(f"trait_contract_synthetic failed ({ERRORS.get(rc, rc)})")
    return z, ms.value


class _Replicate(Simulation):
    """A replicate context borrowed from a Batch (the batch owns and frees it)."""

    def __init__(self, cfg, ctx, owner):
        self.cfg = cfg
        self.ctx = C.c_void_p(ctx)
        self._owner = owner

    def close(self):
        self.ctx = C.c_void_p()


class Batch:
    """Replicate populations advanced together by batched persistent kernels."""

    def __init__(self, cfgs, device: int = 0):
        cfgs = list(cfgs)
        arr = (Config * len(cfgs))()
        for i, c in enumerate(cfgs):
            C.memmove(C.byref(arr[i]), C.byref(c), C.sizeof(Config))
        self.cfgs = cfgs
        self.h = C.c_void_p()
        rc = lib().clotho_batch_create(arr, len(cfgs), device, C.byref(self.h))
        if rc != 0:
            raise ClothoError(f"clotho_batch_create failed ({ERRORS.get(rc, rc)})")
        self._reps = [_Replicate(c, lib().clotho_batch_replicate(self.h, i), self)
                      for i, c in enumerate(cfgs)]

    def __len__(self):
        return len(self._reps)

    def replicate(self, r) -> Simulation:
        return self._reps[r]

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().clotho_batch_last_error(self.h)
            raise ClothoError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def run(self, n=1, mode="graph"):
        """Advance every replicate n generations; returns device milliseconds."""
        ms = C.c_double(0.0)
        self._check(lib().clotho_batch_run(self.h, int(n), BATCH_MODES[mode], C.byref(ms)),
                    f"batch_run({mode})")
        return ms.value

    def autotune(self):
        """Persistent grid size of the batched launches, percent of the default."""
        pct = C.c_uint32(0)
        self._check(lib().clotho_batch_autotune(self.h, C.byref(pct)), "batch_autotune")
        return pct.value

    def set_timing(self, on=True):
        lib().clotho_batch_set_timing(self.h, 1 if on else 0)

    def timings(self):
        out = np.zeros(6, np.float64)
        n = lib().clotho_batch_timings(self.h, out, 6)
        return n, out

    def close(self):
        if self.h:
            for r in self._reps:
                r.close()
            lib().clotho_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

