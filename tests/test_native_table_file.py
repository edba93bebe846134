"""Format-v7 table sets on disk (c8mi_native_table_*): the payload is the
canonical encoding the reference hashes (calculate_proposal_native_hash,
src/kokkos/em/tables/ProposalNativeTable.cpp:570-666), so the file's SHA-256 is
the table's content hash.

* SHA-256 against the reference test's golden vectors
  (tests/kokkos/em/tables/testNativeTableContracts.cpp:18-30);
* the content hash of the air, ice and randomised table sets equals the
  reference's own calculate_proposal_native_hash, compiled in place
  (oracle/_ref); any changed value, cut or identity changes it;
* write -> read returns the same arrays (byte for byte), metadata and hash, and
  the read set answers every native query as the original does;
* a damaged file (flipped byte, truncation, wrong magic, trailing bytes) is
  refused with status 1, a missing one with status 4.
"""
import ctypes
import os

import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables
from c8mi import tables as native


def sha(data: bytes) -> str:
    digest = (ctypes.c_uint8 * 32)()
    buf = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
    c8mi._check(c8mi.library().c8mi_sha256(buf, ctypes.c_uint64(len(data)), digest))
    return bytes(digest).hex()


def test_sha256_golden_vectors():
    assert sha(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert sha(b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert sha(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"
    # message lengths around the block and padding boundaries, against hashlib
    import hashlib
    for n in (55, 56, 63, 64, 65, 119, 120, 127, 128, 1000):
        data = bytes((7 * i + 3) % 256 for i in range(n))
        assert sha(data) == hashlib.sha256(data).hexdigest(), n


TABLES = {"air": lambda: airtables.build_air_tables(),
          "ice": lambda: airtables.build_ice_tables(cut_mev=0.1),
          "random": lambda: native.random_table_set(seed=5, noise=0.01)}


def meta():
    return native.table_meta(stochastic_cut_mev=0.5, identities=[(22, 11, 21), (11, 11, 22), (-11, 11, 23)])


@pytest.mark.parametrize("name", list(TABLES))
def test_content_hash_equals_the_reference(name):
    ref = oracle_lib.load_ref()
    if ref is None:
        pytest.skip("the reference build (oracle/_ref) is not present")
    t, m = TABLES[name](), meta()
    want = (ctypes.c_uint8 * 32)()
    ref.ref_native_table_hash(ctypes.byref(t.struct), ctypes.byref(m), want)
    assert native.table_hash(t, m) == bytes(want)


def test_content_hash_covers_values_metadata_and_identities():
    t, m = TABLES["random"](), meta()
    base = native.table_hash(t, m)
    assert native.table_hash(t, meta()) == base
    t.cubic_values[17] = np.nextafter(t.cubic_values[17], np.inf)
    assert native.table_hash(t, m) != base
    t2 = TABLES["random"]()
    m2 = meta()
    m2.stochastic_cut_mev = 0.4
    assert native.table_hash(t2, m2) != base
    m3 = native.table_meta(stochastic_cut_mev=0.5, identities=[(22, 11, 21), (11, 11, 22), (-11, 11, 24)])
    assert native.table_hash(t2, m3) != base
    m4 = native.table_meta(proposal_version="7.6.1", stochastic_cut_mev=0.5,
                           identities=[(22, 11, 21), (11, 11, 22), (-11, 11, 23)])
    assert native.table_hash(t2, m4) != base


def same_fields(a, b):
    """Field by field, byte for byte: struct padding carries no content."""
    if a.dtype.names is None:
        return a.shape == b.shape and np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()
    return all(same_fields(a[k], b[k]) for k in a.dtype.names)


def _array(pointer, dtype, count):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    raw = (ctypes.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(pointer)
    return np.frombuffer(bytes(raw), dtype=dtype)


@pytest.mark.parametrize("name", ["air", "random"])
def test_round_trip(tmp_path, name):
    t, m = TABLES[name](), meta()
    path = tmp_path / f"{name}.c8mint"
    native.write_table_file(path, t, m)
    assert not [p for p in os.listdir(tmp_path) if ".tmp." in p]
    f = native.TableFile(path)
    assert f.digest == native.table_hash(t, m)
    s = f.struct
    assert (s.dndx_column_count, s.total_rate_column_count, s.dedx_column_count, s.utility_column_count) == \
        (t.dndx.size, t.totals.size, t.dedx.size, t.utility.size)
    assert same_fields(_array(s.dndx_columns, native.DNDX_DTYPE, s.dndx_column_count), t.dndx)
    assert same_fields(_array(s.total_rate_columns, native.TOTAL_DTYPE, s.total_rate_column_count), t.totals)
    assert same_fields(_array(s.dedx_columns, native.DEDX_DTYPE, s.dedx_column_count), t.dedx)
    assert same_fields(_array(s.utility_columns, native.UTILITY_DTYPE, s.utility_column_count), t.utility)
    assert np.array_equal(_array(s.selection_column_indices, np.uint32, s.dndx_column_count), t.selection)
    assert np.array_equal(_array(s.cubic_values, np.float64, s.cubic_coefficient_count), t.cubic_values)
    assert np.array_equal(_array(s.bicubic_mixed_derivative, np.float64, s.bicubic_coefficient_count),
                          t.bicubic_mixed)
    assert f.meta.stochastic_cut_mev == 0.5 and f.meta.proposal_version == b"7.6.2"
    assert f.meta.interaction_identity_count == 3
    for k in ("electron_mass_mev", "muon_mass_mev", "sqrt_e"):
        assert getattr(s.constants, k) == getattr(t.struct.constants, k)
    # the read set answers every query as the original
    queries = native.random_queries(t, 3000, seed=9)
    want = oracle_lib.native_queries(t, queries)
    got = oracle_lib.native_queries(f, queries)
    assert np.array_equal(got["status"], want["status"])
    assert np.array_equal(got["value"], want["value"], equal_nan=True)
    f.close()


def test_damaged_files_are_refused(tmp_path):
    t, m = TABLES["air"](), meta()
    path = tmp_path / "air.c8mint"
    native.write_table_file(path, t, m)
    good = path.read_bytes()
    cases = {"flipped": good[:100] + bytes([good[100] ^ 1]) + good[101:],
             "truncated": good[:-40],
             "magic": b"C8MINT08" + good[8:],
             "digest": good[:-1] + bytes([good[-1] ^ 0x80])}
    for name, data in cases.items():
        bad = tmp_path / f"{name}.c8mint"
        bad.write_bytes(data)
        with pytest.raises(c8mi.C8miError) as err:
            native.TableFile(bad)
        assert err.value.status == 1, name
    with pytest.raises(c8mi.C8miError) as err:
        native.TableFile(tmp_path / "missing.c8mint")
    assert err.value.status == 4
