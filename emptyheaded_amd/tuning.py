"""Committed tuning state: the autotuner's chosen knobs (tuning.json) and the
learned layout selector's weights (selector.json), both produced on an MI355X
by emptyheaded_amd.autotune and emptyheaded_amd.selector."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_JSON = os.path.join(HERE, "tuning.json")
SELECTOR_JSON = os.path.join(HERE, "selector.json")


_selector_active = False


def selector_active():
    """True once apply_committed has loaded selector weights."""
    return _selector_active


_plan_knobs = {}
_tuning_stamp = None


def plan_knobs(plan_name):
    """The committed knobs with the overrides the autotuner found for one plan
    (tuning.json "plans"; unknown plans get the global knobs).  The file is
    parsed once and again only when it changes on disk."""
    global _tuning_stamp
    st = os.stat(TUNING_JSON)
    stamp = (st.st_mtime_ns, st.st_size)
    if stamp != _tuning_stamp:
        _plan_knobs.clear()
        _tuning_stamp = stamp
    if plan_name not in _plan_knobs:
        with open(TUNING_JSON) as f:
            t = json.load(f)
        params = dict(t["best"])
        params.update(t.get("plans", {}).get(plan_name, {}).get("best", {}))
        _plan_knobs[plan_name] = params
    return dict(_plan_knobs[plan_name])


def apply_plan(plan_name):
    """Apply the committed knobs of one plan (plan_knobs) process-wide."""
    import emptyheaded_amd as E
    return E.tune_set(**plan_knobs(plan_name))


def apply_committed(device=0, selector_path=None):
    global _selector_active
    import emptyheaded_amd as E
    applied = {"params": None, "selector": None}
    if os.path.exists(TUNING_JSON):
        with open(TUNING_JSON) as f:
            t = json.load(f)
        applied["params"] = E.tune_set(**t["best"])
    else:
        applied["params"] = E.tune_get()
    sp = selector_path or SELECTOR_JSON
    if os.path.exists(sp):
        from . import selector
        applied["selector"] = selector.load_committed(device, sp)
        _selector_active = bool(applied["selector"])
    from . import pairsel
    if os.path.exists(pairsel.PAIR_JSON):
        applied["pair_selector"] = pairsel.load_committed(device)
    return applied
