"""The build check of the stage, line-product and clover kernels'
global-access addresses (msc_project_amd/check_load_bases.py): planted
defects in small hand-written gfx950 listings are found on exactly the
direction whose path carries them, paths that no wave can take are not
reported, and the compiled library's kernels pass."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "msc_project_amd"))
import check_load_bases as clb  # noqa: E402

NAME = "_ZN3lgf11k_rk3_stageItest"
BASE = 0x1000
TIDY_SETUP = (2, True, False, 0)  # kernarg pointer in s[0:1], blockIdx.x in s2, direction from threadIdx.y
WGID_SETUP = (2, True, True, 0)  # direction from blockIdx.y in s3

PROLOGUE = [
    "s_load_dwordx2 s[4:5], s[0:1], 0x0",  # FieldDesc*
    "v_bfe_u32 v1, v0, 10, 10",
    "v_readfirstlane_b32 s6, v1",  # direction = threadIdx.y
    "s_waitcnt lgkmcnt(0)",
    "s_load_dwordx2 s[24:25], s[4:5], 0x60",  # rcp[3], a double
    "s_load_dwordx2 s[8:9], s[4:5], 0x0",  # links[0]
    "s_waitcnt lgkmcnt(0)",
]


def listing(lines, name=NAME):
    """an llvm-objdump style listing; 'label:' lines name addresses, '-> label' ends a branch"""
    out, addr, labels, body = ["%016x <%s>:" % (BASE, name)], BASE, {}, []
    for ln in lines:
        if ln.endswith(":"):
            labels[ln[:-1]] = addr
        else:
            body.append((addr, ln))
            addr += 8
    for a, ln in body:
        code, _, label = ln.partition(" -> ")
        tail = " <%s+0x%x>" % (name, labels[label] - BASE) if label else ""
        out.append("\t%-58s // %012X: BF800000%s" % (code, a, tail))
    return "\n".join(out) + "\n"


def findings(lines, setup=TIDY_SETUP, name=NAME):
    (instrs,) = clb.parse(listing(lines, name)).values()
    found, checked = clb.check_kernel(instrs, setup, clb.family(name))
    return found, checked


def test_merged_tail_with_stale_base_is_found_on_its_direction_only():
    # direction 3 jumps to the merged block before the base pair is overwritten
    lines = PROLOGUE + [
        "s_cmp_eq_u32 s6, 3",
        "s_cbranch_scc1 0 -> merge",
        "s_add_u32 s24, s8, 0x100",
        "s_addc_u32 s25, s9, 0",
        "merge:",
        "global_load_dwordx4 v[2:5], v6, s[24:25]",
        "s_endpgm",
    ]
    found, checked = findings(lines)
    assert checked == 1
    assert {f[3] for f in found} == {3}
    assert all(("desc", 96) in f[2] or ("desc", 100) in f[2] for f in found)


def test_every_path_setting_the_base_passes():
    lines = PROLOGUE + [
        "s_cmp_eq_u32 s6, 3",
        "s_cbranch_scc0 0 -> other",
        "s_add_u32 s24, s8, 0x200",
        "s_addc_u32 s25, s9, 0",
        "s_branch 0 -> merge",
        "other:",
        "s_add_u32 s24, s8, 0x100",
        "s_addc_u32 s25, s9, 0",
        "merge:",
        "global_load_dwordx4 v[2:5], v6, s[24:25]",
        "s_endpgm",
    ]
    assert findings(lines) == ([], 1)


def test_correlated_branches_leave_no_false_path():
    # directions 0 and 1 skip both the base set-up and the load: the path that
    # sets nothing and loads does not exist, though a path-insensitive reading sees it
    lines = PROLOGUE + [
        "s_cmp_lt_u32 s6, 2",
        "s_cbranch_scc1 0 -> low",
        "s_add_u32 s24, s8, 0x100",
        "s_addc_u32 s25, s9, 0",
        "low:",
        "s_cmp_lt_u32 s6, 2",
        "s_cbranch_scc1 0 -> end",
        "global_load_dwordx4 v[2:5], v6, s[24:25]",
        "end:",
        "s_endpgm",
    ]
    assert findings(lines) == ([], 1)
    # with the second test on another direction set, directions 2.. skip and 0, 1 load: found
    bad = [ln.replace("s_cmp_lt_u32 s6, 2", "s_cmp_gt_u32 s6, 1") if k > len(PROLOGUE) + 3 else ln
           for k, ln in enumerate(lines)]
    found, _ = findings(bad)
    assert {f[3] for f in found} == {0, 1}


def test_direction_from_the_workgroup_id():
    lines = [
        "s_load_dwordx2 s[4:5], s[0:1], 0x0",
        "s_waitcnt lgkmcnt(0)",
        "s_load_dwordx2 s[24:25], s[4:5], 0x60",
        "s_load_dwordx2 s[8:9], s[4:5], 0x8",  # links[1]
        "s_waitcnt lgkmcnt(0)",
        "s_cmp_lg_u32 s3, 2",
        "s_cbranch_scc0 0 -> merge",
        "s_mov_b64 s[24:25], s[8:9]",
        "merge:",
        "global_load_dwordx2 v[2:3], v6, s[24:25]",
        "s_endpgm",
    ]
    found, _ = findings(lines, WGID_SETUP)
    assert {f[3] for f in found} == {2}


def test_exec_zero_paths_do_not_access_memory():
    # the load on the path where no lane is active cannot fault; the one after
    # a branch that leaves lanes active is reported
    lines = PROLOGUE + [
        "v_cmp_gt_i32_e32 vcc, s6, v0",
        "s_and_saveexec_b64 s[10:11], vcc",
        "s_cbranch_execnz 0 -> live",
        "global_load_dwordx4 v[2:5], v6, s[24:25]",
        "s_endpgm",
        "live:",
        "global_load_dwordx4 v[2:5], v6, s[8:9]",
        "s_endpgm",
    ]
    assert findings(lines) == ([], 1)
    lines[-2] = "global_load_dwordx4 v[2:5], v6, s[24:25]"
    found, _ = findings(lines)
    assert {f[0] for f in found} == {BASE + 8 * (len(PROLOGUE) + 5)} and {f[3] for f in found} == {0, 1, 2, 3}


def test_vcc_from_exec_and_a_constant_mask():
    # the compiler's pattern: a case flag in an SGPR pair, tested as exec & ~flag
    lines = PROLOGUE + [
        "s_cmp_lg_u32 s6, 0",
        "s_cselect_b64 s[26:27], -1, 0",  # flag = direction != 0
        "s_andn2_b64 vcc, exec, s[26:27]",
        "s_cbranch_vccz 0 -> skip",  # directions != 0 skip the load
        "global_load_dwordx4 v[2:5], v6, s[24:25]",
        "skip:",
        "s_endpgm",
    ]
    found, _ = findings(lines)
    assert {f[3] for f in found} == {0}


def test_vector_addresses_and_lane_spills():
    lines = PROLOGUE + [
        "v_writelane_b32 v40, s8, 0",  # spill links[0] to a lane
        "v_writelane_b32 v40, s9, 1",
        "s_mov_b64 s[8:9], 0",
        "v_readlane_b32 s12, v40, 0",
        "v_readlane_b32 s13, v40, 1",
        "v_mov_b64_e32 v[10:11], s[12:13]",
        "v_lshl_add_u64 v[10:11], v[6:7], 4, v[10:11]",
        "global_load_dwordx4 v[2:5], v[10:11], off",
        "v_mov_b64_e32 v[12:13], s[24:25]",
        "global_store_dwordx4 v[12:13], v[2:5], off",
        "s_endpgm",
    ]
    found, checked = findings(lines)
    assert checked == 2
    assert {f[0] for f in found} == {BASE + 8 * (len(PROLOGUE) + 9)}


def test_vcc_halves_as_address_registers():
    # the compiler also uses vcc_lo / vcc_hi as two plain SGPRs in address arithmetic
    lines = PROLOGUE + [
        "s_load_dwordx2 s[20:21], s[4:5], 0x38",  # a stride field of the descriptor
        "s_waitcnt lgkmcnt(0)",
        "s_add_u32 vcc_lo, s8, s20",  # links[0] + stride
        "s_addc_u32 vcc_hi, s9, s21",
        "s_add_u32 s12, vcc_lo, s20",
        "s_addc_u32 s13, vcc_hi, s21",
        "global_load_dwordx4 v[2:5], v6, s[12:13]",
        "v_cmp_gt_u32_e32 vcc, 4, v6",  # VCC rewritten as a lane mask
        "s_add_u32 s14, vcc_lo, s20",
        "s_addc_u32 s15, vcc_hi, s21",
        "global_load_dwordx4 v[2:5], v6, s[14:15]",
        "s_endpgm",
    ]
    found, checked = findings(lines)
    assert checked == 2
    assert {f[0] for f in found} == {BASE + 8 * (len(PROLOGUE) + 10)}


def test_pointer_arguments_of_the_clover_kernel():
    # k_clover(fds, tiles, int t0, int ping, lines, int64_t G, partials): bytes 24 and 40 are pointers, 16 and 32 not
    name = "_ZN3lgf8k_clovertest"
    lines = PROLOGUE + [
        "s_load_dwordx2 s[10:11], s[0:1], 0x18",  # lines
        "s_load_dwordx2 s[12:13], s[0:1], 0x28",  # partials
        "s_load_dwordx2 s[14:15], s[0:1], 0x20",  # G
        "s_waitcnt lgkmcnt(0)",
        "global_load_dwordx4 v[2:5], v6, s[10:11]",
        "global_store_dwordx2 v6, v[2:3], s[12:13]",
        "s_cmp_eq_u32 s6, 5",  # only plane 5 adds G to a pointer by mistake
        "s_cselect_b64 s[16:17], s[14:15], s[10:11]",
        "global_load_dwordx4 v[2:5], v6, s[16:17]",
        "s_endpgm",
    ]
    found, checked = findings(lines, name=name)
    assert checked == 3
    assert {f[0] for f in found} == {BASE + 8 * (len(PROLOGUE) + 8)} and {f[3] for f in found} == {5}
    assert clb.family(name) == (6, (0, 8, 24, 40))
    assert clb.family("_ZN3lgf7k_linesItest") == (4, (0, 8, 24))


def test_elided_padding_lines_are_skipped():
    text = listing(PROLOGUE + ["global_load_dwordx4 v[2:5], v6, s[8:9]", "s_endpgm"]) + "\t\t...\n"
    (instrs,) = clb.parse(text).values()
    assert all(isinstance(a, int) for a, _, _, _ in instrs)


OBJ = os.path.join(ROOT, "msc_project_amd", "_build", "kernels.hip.o")


@pytest.mark.skipif(not os.path.exists(OBJ), reason="the kernel object is built by __graft_entry__.build()")
def test_compiled_kernels_pass():
    results = clb.check_object(OBJ)
    assert len(results) >= 40
    assert any(k.startswith("_ZN3lgf7k_lines") for k in results)
    assert any(k.startswith("_ZN3lgf8k_clover") for k in results)
    assert sum(n for _, n in results.values()) > 10000
    bad = [(k, f) for k, (fs, _) in results.items() for f in fs]
    assert not bad, bad[:5]
