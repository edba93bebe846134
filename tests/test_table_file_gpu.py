"""A table set read back from a format-v7 file (c8mi_native_table_read) drives
the device exactly as the set it was written from: the same seeded showers give
identical counters, ledgers and profiles, bit for bit."""
import numpy as np
import pytest

from c8mi import airtables, em
from c8mi import tables as native

pytestmark = pytest.mark.gpu


def shower(table_struct_owner, pid, energy, seed):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True)
    dev = em.EmBackend(cfg, table_struct_owner, atm)
    dev.begin_shower(seed, 2)
    dev.enqueue(em.particle(pid, energy, (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0), history_id=1))
    result = dev.run()
    hist, ledger = dev.download_profile()
    dev.close()
    return result, hist, ledger


def test_showers_from_a_table_file_equal_the_original(tmp_path):
    tables = airtables.build_air_tables()
    path = tmp_path / "air.c8mint"
    meta = native.table_meta(stochastic_cut_mev=0.5)
    native.write_table_file(path, tables, meta)
    read = native.TableFile(path)
    assert read.digest == native.table_hash(tables, meta)
    for pid, energy, seed in ((22, 10.0, 1), (11, 3.0, 2)):
        print(f"table file shower: pid {pid}, {energy} GeV, seed {seed}", flush=True)
        r0, h0, l0 = shower(tables, pid, energy, seed)
        r1, h1, l1 = shower(read, pid, energy, seed)
        assert r0 == r1
        assert l0 == l1
        for k in h0:
            assert np.array_equal(h0[k], h1[k]), k
    read.close()
