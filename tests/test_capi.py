"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public headers declare, and its host-side validation and table arithmetic
behave like the reference's (RadioAccumulator.inl:934-975,
FlatAtmosphereConfig.inl:31-73)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import c8mi
import oracle_lib
from radio_cases import case_observers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "c8mi", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for match in re.finditer(r"\b(c8mi_[a-z0-9_]+)\s*\(", text):
            names.add(match.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(c8mi_lib):
    names = declared_symbols()
    assert "c8mi_radio_create" in names
    missing = [n for n in names if not hasattr(c8mi_lib, n)]
    assert not missing, missing


def test_flat_table_matches_oracle_bitwise():
    refractivity = 3.25e-4 * np.exp(-(np.arange(20001) - 1000.0) / 8000.0)
    prop = c8mi.Propagation(refractivity=refractivity, minimum_height_m=-1000.0,
                            maximum_height_m=19000.0, step_m=1.0)
    integrated, slopes = oracle_lib.build_flat_table(refractivity, 1.0)
    assert np.array_equal(prop.integrated, integrated)
    s = prop.struct
    assert (s.slope_refractivity_lower, s.slope_integrated_refractivity_lower,
            s.slope_refractivity_upper, s.slope_integrated_refractivity_upper) == tuple(slopes)
    assert s.table_size == refractivity.size and s.inverse_step_per_m == 1.0


def test_flat_table_rejects_short_tables():
    with pytest.raises(c8mi.C8miError) as err:
        c8mi.Propagation(refractivity=np.zeros(10), minimum_height_m=0.0, maximum_height_m=9.0)
    assert err.value.status == 1


def _bad_config(**changes):
    obs = case_observers(2e-6)
    prop = c8mi.Propagation(refractivity=np.full(3001, 3e-4), minimum_height_m=-1000.0,
                            maximum_height_m=2000.0)
    kwargs = dict(observers=obs, propagation=prop)
    kwargs.update(changes)
    return kwargs


@pytest.mark.parametrize("changes", [
    dict(coreas=False, zhs=False),
    dict(fixed_point_limit=0.0),
    dict(fixed_point_limit=float("nan")),
    dict(zhs_refinement=0),
    dict(zhs_refinement=65),
])
def test_invalid_configs_are_rejected_before_device_use(changes):
    with pytest.raises(c8mi.C8miError) as err:
        c8mi.RadioAccumulator(**_bad_config(**changes))
    assert err.value.status == 1
    assert c8mi.library().c8mi_last_error()


def test_observer_outside_table_is_rejected():
    obs = case_observers(2e-6, positions=((0.0, 0.0, 5000.0),))
    with pytest.raises(c8mi.C8miError) as err:
        c8mi.RadioAccumulator(**_bad_config(observers=obs))
    assert err.value.status == 1 and "outside" in err.value.message


@pytest.mark.parametrize("field,value", [("sample_rate_hz", 0.0), ("number_of_bins", 0),
                                         ("duration_s", -1.0), ("start_time_s", float("inf"))])
def test_invalid_observer_is_rejected(field, value):
    obs = case_observers(2e-6)
    obs[1][field] = value
    with pytest.raises(c8mi.C8miError) as err:
        c8mi.RadioAccumulator(**_bad_config(observers=obs))
    assert err.value.status == 1


def test_record_dtype_matches_c_layout():
    assert c8mi.PARTICLE_STATE_DTYPE.itemsize == 128
    assert c8mi.LEPTON_TRANSPORT_RECORD_DTYPE.itemsize == 752
    assert c8mi.LEPTON_TRANSPORT_RECORD_DTYPE.fields["start"][1] == 320
    assert c8mi.LEPTON_TRANSPORT_RECORD_DTYPE.fields["end"][1] == 448
    assert ctypes.sizeof(c8mi.RadioObserver) == c8mi.OBSERVER_DTYPE.itemsize == 56


# ---- EM tables: bounds validation runs before any device call ---------------

def _bicubic_column(t):
    from c8mi import tables
    return int(np.flatnonzero(t.dndx["rate_model"] == tables.RATE_BICUBIC)[0])


def _break_bicubic_offset(t):
    t.dndx["spline"]["coefficient_offset"][_bicubic_column(t)] = t.bicubic_values.size


def _break_bicubic_rows(t):
    t.dndx["spline"]["rows"][_bicubic_column(t)] += 1


def _break_selection(t):
    t.selection[0] = t.dndx.size


def _break_cubic(t):
    t.totals["spline"]["coefficient_count"][0] = t.cubic_values.size + 1


def _break_order(t):
    t.dndx[[0, 1]] = t.dndx[[1, 0]]


def _break_dedx_index(t):
    t.dedx_order[0] = t.dedx.size


def _break_polynomial(t):
    i = int(np.flatnonzero(t.dndx["spline"]["polynomial_coefficient_offset"] != 2**64 - 1)[0])
    t.dndx["spline"]["polynomial_coefficient_offset"][i] = t.polynomial.size - 8


@pytest.mark.parametrize("breaker", [_break_bicubic_offset, _break_bicubic_rows, _break_selection,
                                     _break_cubic, _break_order, _break_dedx_index,
                                     _break_polynomial])
def test_em_tables_reject_out_of_bounds_descriptors(c8mi_lib, breaker):
    import c8mi
    from c8mi import tables

    t = tables.random_table_set(1)
    breaker(t)
    with pytest.raises(c8mi.C8miError) as err:
        c8mi.EmTables(t)
    assert err.value.status == 1, err.value.message
    assert "native tables" in err.value.message


def test_random_key_dtype_matches_c_layout():
    import c8mi

    assert c8mi.RANDOM_KEY_DTYPE.itemsize == 48
    assert c8mi.RANDOM_KEY_DTYPE.fields["draw_id"][1] == 40


def test_em_binding_structs_match_c_layout(tmp_path):
    """The ctypes mirrors of the EM ABI structs have the C sizes (gcc on the
    public header)."""
    from c8mi import em

    names = {"c8mi_em_config": em.EmConfig, "c8mi_em_run_result": em.RunResult,
             "c8mi_em_ledger": em.Ledger, "c8mi_em_shower_axis": em.ShowerAxis,
             "c8mi_em_slot_state": em.SlotState, "c8mi_em_request": em.Request,
             "c8mi_em_serve_options": em.ServeOptions, "c8mi_em_request_result": em.RequestResult,
             "c8mi_environment": em.Environment, "c8mi_env_node": em.EnvNode,
             "c8mi_env_model": em.EnvModel, "c8mi_em_wavefront_result": em.WavefrontResult,
             "c8mi_em_limits": em.Limits, "c8mi_em_statistics": em.Statistics}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "c8mi/c8mi_em.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({n}));\n' for n in names) + "return 0;}\n")
    exe = tmp_path / "sizes"
    import subprocess

    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                            text=True).stdout.split()]
    for (name, cls), size in zip(names.items(), sizes):
        assert ctypes.sizeof(cls) == size, name
    assert em.FALLBACK_DTYPE.itemsize == 144
