"""Densification parameters (``pmp_params`` of include/pmp/pmp.h) as ctypes.

Defaults are the reference's operating values:
  patch half 5 at stride 2 (``depthmap.cc:666,725`` patch_size_-2 with
  patch_size 7, ``config.cc:183``; taps ``PatchMatchCuda.cu:104-106``),
  sigma_color 50 / sigma_space 5 (``PatchMatchCuda.cu:897-898``),
  variance floor 0.1 and fold agreement 0.25 (``:170-171``),
  penalty gains 1250 / 8 (``:326,330``), refinement 0.02 / 0.5 with decays
  0.3 / 0.8 over 6 trials (``:348-349,364,393-394``), min score 0.1
  (``depthmap_computation_module.cc:223``), depth range [1e-5, ...]
  (``depthmap_computation_module.cc:123-131``), 2 iterations
  (``config.cc`` default ``Dense.Iterations``), median test 0.05
  (``depthmap.cc:1190``, a double: the float ratio is compared with the
  double literal there).
"""
import ctypes

# REF_FOLD: the fold of PatchMatchCuda.cu:161-171 over moments pooled across
# views (zeroed once at :57-62, outside the view loop :101); FOLD_INDEP: the
# same fold over independent per-view NCCs (no counterpart in the reference)
AGG_REF_FOLD, AGG_MAX, AGG_TOPK_MEAN, AGG_FOLD_INDEP = 0, 1, 2, 3
EDGE_WRAP, EDGE_CLAMP, EDGE_ZERO = 0, 1, 2
COVER_REF_GRID, COVER_FULL = 0, 1
INIT_REF_UNIFORM, INIT_INVDEPTH_HEMI = 0, 1


class Params(ctypes.Structure):
    _fields_ = [
        ("patch_half", ctypes.c_int32), ("patch_stride", ctypes.c_int32),
        ("iterations", ctypes.c_int32), ("n_refine", ctypes.c_int32),
        ("agg_mode", ctypes.c_int32), ("topk", ctypes.c_int32),
        ("edge_mode", ctypes.c_int32), ("coverage_mode", ctypes.c_int32),
        ("init_mode", ctypes.c_int32), ("postprocess", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("min_depth", ctypes.c_float), ("max_depth", ctypes.c_float),
        ("init_inv_depth_min", ctypes.c_float), ("init_inv_depth_max", ctypes.c_float),
        ("min_score", ctypes.c_float), ("sigma_color", ctypes.c_float),
        ("sigma_space", ctypes.c_float), ("var_floor", ctypes.c_float),
        ("fold_agree", ctypes.c_float),
        ("depth_penalty", ctypes.c_float), ("plane_penalty", ctypes.c_float),
        ("refine_depth_range", ctypes.c_float), ("refine_normal_range", ctypes.c_float),
        ("refine_depth_decay", ctypes.c_float), ("refine_normal_decay", ctypes.c_float),
        ("min_patch_sd", ctypes.c_float), ("median_rel_thr", ctypes.c_double),
    ]

    def copy(self, **kw):
        p = Params()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def default_params(**kw) -> Params:
    """Reference defaults (see module docstring); keyword overrides."""
    p = Params(
        patch_half=5, patch_stride=2, iterations=2, n_refine=6,
        agg_mode=AGG_REF_FOLD, topk=2, edge_mode=EDGE_WRAP, coverage_mode=COVER_REF_GRID,
        init_mode=INIT_REF_UNIFORM, postprocess=1, seed=0,
        min_depth=1e-5, max_depth=10.0, init_inv_depth_min=1.0 / 20.0, init_inv_depth_max=1.0 / 0.3,
        min_score=0.1, sigma_color=50.0, sigma_space=5.0, var_floor=0.1, fold_agree=0.25,
        depth_penalty=1250.0, plane_penalty=8.0,
        refine_depth_range=0.02, refine_normal_range=0.5, refine_depth_decay=0.3, refine_normal_decay=0.8,
        min_patch_sd=0.0, median_rel_thr=0.05,
    )
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def config0_params(**kw) -> Params:
    """Config 0: 7x7 patch at stride 1, inverse-depth/hemisphere init, 4 iterations."""
    d = dict(patch_half=3, patch_stride=1, iterations=4, init_mode=INIT_INVDEPTH_HEMI,
             max_depth=20.0, seed=0)
    d.update(kw)
    return default_params(**d)


def config1_params(**kw) -> Params:
    """Config 1: 11x11 patch at stride 2, best-3-of-6 mean, 8 iterations, full coverage."""
    d = dict(patch_half=5, patch_stride=2, iterations=8, agg_mode=AGG_TOPK_MEAN, topk=3,
             init_mode=INIT_INVDEPTH_HEMI, coverage_mode=COVER_FULL, max_depth=20.0, seed=0)
    d.update(kw)
    return default_params(**d)


# ---- JSON configs for the tools (bench.py --params, profile_run --params)

CONFIGS = {"default": default_params, "config0": config0_params, "config1": config1_params}
ENUMS = {
    "agg_mode": {"ref_fold": AGG_REF_FOLD, "max": AGG_MAX, "topk_mean": AGG_TOPK_MEAN, "fold_indep": AGG_FOLD_INDEP},
    "edge_mode": {"wrap": EDGE_WRAP, "clamp": EDGE_CLAMP, "zero": EDGE_ZERO},
    "coverage_mode": {"ref_grid": COVER_REF_GRID, "full": COVER_FULL},
    "init_mode": {"ref_uniform": INIT_REF_UNIFORM, "invdepth_hemi": INIT_INVDEPTH_HEMI},
}


def params_from_json(src) -> Params:
    """Params from a JSON config: a path, a JSON text or a dict such as
    {"base": "config1", "iterations": 4, "agg_mode": "max"}.  "base" names
    the starting values (default, config0, config1; default config1); every
    other key is a pmp_params field, with enum fields given by value or by
    name.  Unknown keys and values of the wrong type raise ValueError (the
    library checks ranges when the params are used)."""
    import json
    import os
    if isinstance(src, dict):
        d = dict(src)
    elif isinstance(src, str) and src.lstrip().startswith(("{", "[")):
        d = json.loads(src)
    else:
        if not os.path.exists(src):
            raise ValueError(f"no such params file: {src}")
        with open(src) as f:
            d = json.load(f)
    if not isinstance(d, dict):
        raise ValueError("a params config is a JSON object")
    base = d.pop("base", "config1")
    if base not in CONFIGS:
        raise ValueError(f"unknown base {base!r} (one of {sorted(CONFIGS)})")
    types = dict(Params._fields_)
    kw = {}
    for k, v in d.items():
        if k not in types:
            raise ValueError(f"unknown params field {k!r}")
        if k in ENUMS and isinstance(v, str):
            if v not in ENUMS[k]:
                raise ValueError(f"{k}: unknown value {v!r} (one of {sorted(ENUMS[k])})")
            v = ENUMS[k][v]
        floating = types[k] in (ctypes.c_float, ctypes.c_double)
        if isinstance(v, bool) or not isinstance(v, (int, float)) or (not floating and not isinstance(v, int)):
            raise ValueError(f"{k}: {'a number' if floating else 'an integer'} expected, got {v!r}")
        kw[k] = v
    return CONFIGS[base](**kw)


def params_to_json(p: Params) -> str:
    """Every field of p as a JSON object (base "default", so that
    params_from_json gives p back)."""
    import json
    return json.dumps({"base": "default", **p.as_dict()}, indent=1)
