"""Host-side marshalling of the drop-in call (gatk_amd/pairhmm.py): the planes handed to the C-ABI
hold every item's bytes in order, whatever form the items come in, and the holders' pointers and
lengths address them (VectorLoglessPairHMM passes Java byte[] arrays; Python callers may pass
uint8 arrays, bytes, lists, other integer dtypes or strided views)."""
import ctypes

import numpy as np
import pytest

from gatk_amd.pairhmm import (READ_HOLDER_DTYPE, ReadDataHolder, _as_u8, _concat_u8, marshal_planes,
                              marshal_reads)


def _expected(items):
    return np.concatenate([_as_u8(x) for x in items])


@pytest.mark.parametrize("kind", ["u8", "bytes", "mixed", "int64", "list", "strided", "str", "int8"])
def test_concat_matches_per_item_conversion(kind):
    rng = np.random.default_rng(1)
    base = [rng.integers(0, 256, int(n)).astype(np.uint8) for n in rng.integers(1, 40, 12)]
    if kind == "u8":
        items = base
    elif kind == "bytes":
        items = [bytes(x) for x in base]
    elif kind == "mixed":
        items = [bytes(x) if k % 2 else x for k, x in enumerate(base)]
    elif kind == "int64":
        items = [x.astype(np.int64) for x in base]
    elif kind == "list":
        items = [list(map(int, x)) for x in base]
    elif kind == "strided":
        items = [np.repeat(x, 2)[::2] for x in base]
    elif kind == "str":
        items = ["".join(chr(65 + int(v) % 26) for v in x) for x in base]
    else:
        items = [x.astype(np.int8) for x in base]
    n_bytes = sum(len(x) for x in items)
    got = _concat_u8(items, n_bytes)
    want = _expected(items)
    assert got.dtype == np.uint8 and got.ndim == 1
    assert np.array_equal(got, want)


def test_holders_address_each_items_bytes():
    rng = np.random.default_rng(2)
    reads = [ReadDataHolder(*(rng.integers(0, 256, n).astype(np.uint8) for _ in range(5))) for n in (1, 7, 150, 33)]
    holders, keep = marshal_reads(reads)
    assert list(holders["length"]) == [1, 7, 150, 33]
    fields = ("read_bases", "read_quals", "insertion_gop", "deletion_gop", "overall_gcp")
    for i, r in enumerate(reads):
        for name, attr in zip(READ_HOLDER_DTYPE.names[:5], fields):
            ptr, n = int(holders[name][i]), int(holders["length"][i])
            got = np.ctypeslib.as_array((ctypes.c_uint8 * n).from_address(ptr))
            assert np.array_equal(got, getattr(r, attr)), (i, name)


def test_unequal_lengths_are_rejected():
    a = np.zeros(5, np.uint8)
    with pytest.raises(ValueError):
        marshal_planes(READ_HOLDER_DTYPE, [[a, a], [a, np.zeros(4, np.uint8)], [a, a], [a, a], [a, a]])
