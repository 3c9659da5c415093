"""The build guard against LDS stores over 8 bytes whose data VGPRs a VALU instruction overwrites
within two instructions (gatk_amd/csrc/check_isa.py), on hand-written assembly text.

The guard runs in the build container only; where its file is not present the test skips.
"""
import importlib.util
import os

import pytest

GUARD = os.path.join(os.path.dirname(__file__), "..", "gatk_amd", "csrc", "check_isa.py")


def _guard():
    if not os.path.exists(GUARD):
        pytest.skip("check_isa.py is not present here")
    spec = importlib.util.spec_from_file_location("check_isa", GUARD)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _asm(*lines):
    return "".join(f"\t{l}\n" for l in lines)


def test_overwrite_one_and_two_instructions_after_a_b128_store_is_flagged():
    g = _guard()
    one = _asm("ds_write_b128 v4, v[22:25] offset:400", "v_mul_f64 v[22:23], v[46:47], v[26:27]")
    two = _asm("ds_write_b128 v4, v[22:25] offset:400", "s_nop 0", "v_cndmask_b32_e64 v25, v1, v2, s[2:3]")
    assert len(g.wide_lds_store_overwrites(one)) == 1
    assert len(g.wide_lds_store_overwrites(two)) == 1
    b96 = _asm("ds_write_b96 v4, v[8:10]", "v_add_f32_e32 v10, v1, v2")
    assert len(g.wide_lds_store_overwrites(b96)) == 1


def test_safe_patterns_are_not_flagged():
    g = _guard()
    three_later = _asm("ds_write_b128 v4, v[22:25]", "v_add_f32_e32 v1, v2, v3", "s_nop 1",
                       "v_mul_f64 v[22:23], v[46:47], v[26:27]")
    other_regs = _asm("ds_write_b128 v4, v[22:25]", "v_mul_f64 v[26:27], v[46:47], v[28:29]")
    address_only = _asm("ds_write_b128 v4, v[22:25]", "v_add_u32_e32 v4, 64, v4")
    narrow = _asm("ds_write2_b64 v97, v[28:29], v[30:31] offset1:1",
                  "v_mov_b32_dpp v28, v30 wave_shr:1 row_mask:0xf bank_mask:0xf")
    for text in (three_later, other_regs, address_only, narrow):
        assert g.wide_lds_store_overwrites(text) == []


def test_comments_and_directives_are_ignored():
    g = _guard()
    text = ("\t.p2align 8\n"
            "\tds_write_b128 v4, v[22:25] ; store the ring pair\n"
            "; v_mul_f64 v[22:23], v[0:1], v[2:3]\n"
            "\tv_add_f32_e32 v1, v2, v3 ; v22 mentioned in a comment only\n")
    assert g.wide_lds_store_overwrites(text) == []
