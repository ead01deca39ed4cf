"""Build check of the compiled RK3 stage, line-product and clover kernels
(k_clover, k_clover_pairs, k_clover_combine): every global memory access
must take its address from a pointer.

A shared-backward-link stage variant once faulted the GPU because the
compiler merged the tails of the four direction cases of a `switch` into one
block, and on one direction's path the scalar register pair that the merged
block used as a load base still held a double loaded from the field
descriptor (its rcp[3] field), not a link-plane address.  This check reads
the gfx950 disassembly of the kernels built here (the code object embedded
in _build/kernels.hip.o) and replays that path-by-path reading for every
stage kernel and every direction a wave can take:

- the direction (threadIdx.y or blockIdx.y: 0..3 for the stage and
  line-product kernels, the plane 0..5 for the clover, the plane of a pair
  0..1 for k_clover_pairs, whose pair comes from blockIdx.x and is not
  fixed) is fixed per run, and
  scalar constants, SCC, VCC and EXEC (zero / non-zero) are propagated, so
  that branches on the direction are decided and only the paths that a wave
  of that direction can take are followed (conditional constant
  propagation; states with different EXEC are not merged);
- every register (scalar, vector, lane and scratch spill slot) carries the
  set of possible provenances of its value: which kernel argument or
  field-descriptor field it came from, through which arithmetic;
- a `global_*` or `flat_*` access whose address (scalar base pair, or 64-bit
  vector address) can hold, on a path with active lanes, a value with no
  pointer in its provenance (a descriptor field past the four pointers, an
  integer from the tile list, loaded data) is reported.

    python check_load_bases.py [_build/kernels.hip.o]     exit 1 on findings

The stage kernels take (FieldDesc* fds, int2* tiles, int ping): kernel
argument bytes 0..15 are the two pointers.  The line-product kernel takes
(fds, tiles, int t0, int ping, C* lines, int64_t G) and the clover kernel
(fds, tiles, t0, ping, lines, G, double* partials): pointers at bytes 0, 8,
24 and 40; k_clover_pairs (fds, tiles, t0, ping, C* lines, G) at 0, 8 and 24;
k_clover_combine (fds, tiles, t0, const C* lines, G, double* partials) at 0,
8, 24 and 40 (t0 is followed by 4 bytes of padding).  FieldDesc's pointer fields (links[0], links[1], acc, qx) are its
first 32 bytes.
"""
import concurrent.futures
import os
import re
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
LLVM = "/opt/rocm/llvm/bin"
# (name pattern, directions a wave can take, byte offsets of the kernel's pointer arguments)
FAMILIES = (
    (re.compile(r"_ZN3lgf1[0-9]k_rk3_stage"), 4, (0, 8)),  # k_rk3_stage, k_rk3_stage_pl, k_rk3_stage_pb
    (re.compile(r"_ZN3lgf7k_lines"), 4, (0, 8, 24)),
    (re.compile(r"_ZN3lgf8k_clover"), 6, (0, 8, 24, 40)),
    (re.compile(r"_ZN3lgf14k_clover_pairs"), 2, (0, 8, 24)),  # the plane also depends on blockIdx.x
    (re.compile(r"_ZN3lgf16k_clover_combine"), 1, (0, 8, 24, 40)),
)
STAGE_FAMILY = FAMILIES[0][1:]
# dword offsets of the pointer arguments of the kernel being checked (set by check_kernel)
pointer_dwords = frozenset([0, 4, 8, 12])
MAX_VALUES = 48  # distinct provenances kept per register before collapsing to "unknown"

UNKNOWN = frozenset([("unknown",)])  # the analysis gave up: taken as a possible pointer
LANE = frozenset([("lane",)])  # per-lane arithmetic with no pointer in it
MEM = frozenset([("mem",)])  # data loaded by a vector memory instruction
TID = frozenset([("tid",)])  # packed work-item id (v0 at entry)
TIDY = frozenset([("tidy",)])  # threadIdx.y
IMM = frozenset([("imm",)])  # made from constants only, not evaluated
VCC_HALVES = ("vcc_lo", "vcc_hi")  # used as plain SGPRs in address arithmetic as well as the VCC mask
Z, NZ, U, ONES = "zero", "nonzero", "unknown", "ones"  # EXEC / VCC / mask states
M32 = 0xFFFFFFFF

# scalar instructions that leave SCC alone (all others are taken to write it)
SCC_KEEP = ("s_mov", "s_cmov", "s_cselect", "s_load", "s_buffer_load", "s_waitcnt", "s_nop", "s_branch",
            "s_cbranch", "s_getpc", "s_setpc", "s_swappc", "s_movk", "s_mul_i32", "s_mul_hi", "s_mulk", "s_brev",
            "s_sleep", "s_setprio", "s_barrier", "s_endpgm", "s_icache", "s_memtime", "s_memrealtime",
            "s_getreg", "s_setreg", "s_sendmsg", "s_sext", "s_ff1", "s_flbit", "s_bitset", "s_clause", "s_delay",
            "s_set_gpr_idx", "s_trap", "s_ttrace", "s_denorm", "s_round", "s_wakeup", "s_pack")
NO_DEST = ("s_cmp", "s_bitcmp", "s_branch", "s_cbranch", "s_waitcnt", "s_nop", "s_barrier", "s_endpgm", "s_sendmsg",
           "s_setprio", "s_sleep", "s_trap", "s_setreg", "s_icache", "s_set_gpr_idx", "s_ttrace", "s_denorm",
           "s_round", "s_setkill", "s_wakeup", "s_clause", "s_delay")
PLAIN_ATOMS = ("lane", "mem", "c", "imm", "tid", "tidy")


def sreg_range(tok):
    """register numbers of an SGPR operand ('s5', 's[4:7]'), else None"""
    m = re.fullmatch(r"s(\d+)", tok)
    if m:
        return [int(m.group(1))]
    m = re.fullmatch(r"s\[(\d+):(\d+)\]", tok)
    if m:
        return list(range(int(m.group(1)), int(m.group(2)) + 1))
    return None


def vreg_range(tok):
    """keys of a VGPR operand ('v5', 'v[4:7]'), else None"""
    m = re.fullmatch(r"v(\d+)", tok)
    if m:
        return [("v", int(m.group(1)))]
    m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return [("v", k) for k in range(int(m.group(1)), int(m.group(2)) + 1)]
    return None


def regs(tok):
    return sreg_range(tok) or vreg_range(tok)


def imm(tok):
    """value of an integer literal operand, else None"""
    if re.fullmatch(r"-?(0x[0-9a-fA-F]+|\d+)", tok):
        return int(tok, 0)
    return None


def parse(text):
    """{kernel: [(addr, mnemonic, operands, branch_target_or_None)]}"""
    kernels, cur, name, start = {}, None, None, 0
    for line in text.splitlines():
        m = re.match(r"^([0-9a-f]+) <(\S+)>:$", line)
        if m:
            name, start = m.group(2), int(m.group(1), 16)
            cur = kernels.setdefault(name, []) if family(name) else None
            continue
        if cur is None or not line.startswith("\t"):
            continue
        code, _, comment = line.strip().partition("//")
        parts = code.strip().split(None, 1)
        am = re.match(r"\s*([0-9A-Fa-f]+):", comment)
        if not parts or not am:  # '...' marks elided zero padding
            continue
        mn = parts[0]
        ops = [o.strip() for o in parts[1].split(",")] if len(parts) > 1 else []
        ops = [t for o in ops for t in (o.split() if " " in o and not o.startswith("offset") else [o])]
        tgt = None
        tm = re.search(r"<(\S+?)\+0x([0-9a-f]+)>", comment)
        if tm and mn.startswith(("s_branch", "s_cbranch")):
            tgt = start + int(tm.group(2), 16) if tm.group(1) == name else None
        cur.append((int(am.group(1), 16), mn, ops, tgt))
    return kernels


def blocks_of(instrs):
    """basic blocks as index ranges, and the block of each block-leading address"""
    index = {ins[0]: i for i, ins in enumerate(instrs)}
    leaders = {0}
    for i, (_, mn, _, tgt) in enumerate(instrs):
        if mn.startswith(("s_branch", "s_cbranch", "s_endpgm", "s_setpc")):
            if i + 1 < len(instrs):
                leaders.add(i + 1)
            if tgt is not None and tgt in index:
                leaders.add(index[tgt])
    starts = sorted(leaders)
    blocks = [(s, (starts[k + 1] if k + 1 < len(starts) else len(instrs))) for k, s in enumerate(starts)]
    return blocks, {instrs[s][0]: k for k, s in enumerate(starts)}


def const(x):
    return {frozenset([("c", x & M32)])}


def const64(x):
    """the two dword constants of a 64-bit inline or literal operand (sign-extended)"""
    return [const(x), const(-1 if x < 0 else 0)]


def const_of(vals):
    """the 32-bit constant a register holds, else None"""
    if len(vals) == 1:
        (v,) = vals
        if len(v) == 1:
            (a,) = v
            if a[0] == "c":
                return a[1]
    return None


def combine(sets):
    """provenance values of an arithmetic result: one per combination of the sources' values"""
    out = {frozenset()}
    for vs in sets:
        vs = {frozenset(a for a in v if a[0] != "c") for v in vs}
        out = {a | b for a in out for b in vs}
        if len(out) > MAX_VALUES:
            return {UNKNOWN}
    return {v if v else IMM for v in out}


def load_atom(base_values, off):
    """provenance atom of a scalar load at `off` from a base with these values"""
    atoms = set().union(*base_values)
    if atoms == {("kernarg",)}:
        return ("arg", off)
    if ("arg", 0) in atoms:
        return ("desc", off)
    if ("arg", 8) in atoms:
        return ("tile", off)
    if any(a[0] == "desc" and a[1] < 32 for a in atoms):
        return ("data", off)
    return ("other", off)


def family(name):
    """(directions, pointer argument offsets) of a kernel name, or None"""
    for pat, n, ptrs in FAMILIES:
        if pat.match(name):
            return n, ptrs
    return None


def is_pointer(atom):
    """both dwords of the kernel's pointer arguments and of the descriptor's four pointer fields"""
    return atom == ("unknown",) or (atom[0] == "arg" and atom[1] in pointer_dwords) or \
        (atom[0] == "desc" and atom[1] < 32)


def signed(x):
    return x - (1 << 32) if x & 0x80000000 else x


def eval_scalar(mn, a, b):
    """(result, scc) of a 32-bit scalar ALU op on constants, or None"""
    op = mn[2:].rsplit("_", 1)[0]
    if op == "add":
        r = a + b
        return r & M32, (int(r > M32) if mn.endswith("u32") else None)
    if op == "sub":
        return (a - b) & M32, (int(b > a) if mn.endswith("u32") else None)
    if op == "mul":
        return (a * b) & M32, None
    if op in ("min", "max"):
        f = min if op == "min" else max
        r = f(signed(a), signed(b)) if mn.endswith("i32") else f(a, b)
        return r & M32, int(r == (a if not mn.endswith("i32") else signed(a)))
    if op == "and":
        r = a & b
    elif op == "or":
        r = a | b
    elif op == "xor":
        r = a ^ b
    elif op == "andn2":
        r = a & ~b & M32
    elif op == "lshl":
        r = (a << (b & 31)) & M32
    elif op == "lshr":
        r = a >> (b & 31)
    elif op == "ashr":
        r = (signed(a) >> (b & 31)) & M32
    else:
        return None
    return r, int(r != 0)


def eval_cmp(mn, a, b):
    """SCC of s_cmp_* / s_bitcmp* on constants"""
    if mn.startswith("s_bitcmp"):
        bit = (a >> (b & 31)) & 1
        return int(bit == (1 if mn.startswith("s_bitcmp1") else 0))
    _, _, rel, ty = mn.split("_")[:4]
    if ty.startswith("i"):
        a, b = signed(a), signed(b)
    return int({"eq": a == b, "lg": a != b, "lt": a < b, "le": a <= b, "gt": a > b, "ge": a >= b}[rel])


class State:
    """registers -> sets of provenance values; SCC (0, 1, None); VCC and EXEC (Z, NZ, U)"""

    __slots__ = ("r", "scc", "vcc", "exec")

    def __init__(self, r, scc=None, vcc=U, exec_=NZ):
        self.r, self.scc, self.vcc, self.exec = r, scc, vcc, exec_

    def copy(self):
        return State(dict(self.r), self.scc, self.vcc, self.exec)

    def get(self, k):
        v = self.r.get(k)
        if v is None:
            return {LANE} if k in VCC_HALVES else {frozenset([("entry", str(k))])}
        return v

    def val(self, tok):
        """provenance set of one operand: register(s) or literal"""
        rr = regs(tok)
        if rr:
            return self.get(rr[0]) if len(rr) == 1 else set().union(*[self.get(k) for k in rr])
        if tok in VCC_HALVES:
            return self.get(tok)
        if tok == "vcc":
            return self.get(VCC_HALVES[0]) | self.get(VCC_HALVES[1])
        c = imm(tok)
        if c is not None:
            return const(c)
        return {LANE}  # float literal, exec, m0 ...

    def join(self, o):
        """merge o into self; True if anything changed"""
        changed = False
        for k, v in o.r.items():
            cur = self.r.get(k)
            if cur is None:
                cur = self.get(k)
                new = cur | v
            elif v <= cur:
                continue
            else:
                new = cur | v
            self.r[k] = {UNKNOWN} if len(new) > MAX_VALUES else new
            changed = True
        for k in list(self.r):
            if k not in o.r:
                new = self.r[k] | o.get(k)
                if new != self.r[k]:
                    self.r[k] = {UNKNOWN} if len(new) > MAX_VALUES else new
                    changed = True
        if self.scc != o.scc and self.scc is not None:
            self.scc, changed = None, True
        if self.vcc != o.vcc and self.vcc != U:
            self.vcc, changed = U, True
        return changed


def mask_state(st, tok):
    """Z / NZ / ONES / U of a 64-bit mask operand"""
    if tok == "exec":
        return st.exec
    if tok == "vcc":
        return st.vcc
    c = imm(tok)
    if c is not None:
        return Z if c == 0 else (ONES if c == -1 else NZ)
    rr = sreg_range(tok)
    if rr and len(rr) == 2:
        lo, hi = const_of(st.get(rr[0])), const_of(st.get(rr[1]))
        if lo is not None and hi is not None:
            return Z if lo == 0 and hi == 0 else (ONES if lo == M32 and hi == M32 else NZ)
    return U


def mask_and(a, b):
    if Z in (a, b):
        return Z
    if a == ONES:
        return b
    if b == ONES:
        return a
    return U


def mask_not(a):
    return {Z: ONES, ONES: Z}.get(a, U)


def address_operands(mn, ops):
    """the register operands that form the address of a global_* / flat_* instruction"""
    if mn.startswith("flat_"):
        rr = [o for o in ops if regs(o)]
        return [rr[1] if "load" in mn or (len(rr) == 3) else rr[0]]
    k = next((i for i, o in enumerate(ops) if o == "off" or (sreg_range(o) and len(sreg_range(o)) == 2)), None)
    if k is None:
        return []
    if ops[k] != "off":
        return [ops[k]]
    before = ops[:k]
    if "load" in mn:
        return [before[1]]
    if "store" in mn:
        return [before[0]]
    return [before[0] if len(before) == 2 else before[1]]  # atomics: (addr, data) or (dst, addr, data)


def transfer(st, ins, direction):
    """apply one instruction to st; returns [(address, text, provenance)] of bad memory addresses"""
    _, mn, ops, _ = ins
    findings = []
    if not ops:
        return findings
    if mn.startswith(("global_", "flat_")):
        if st.exec != Z:
            for tok in address_operands(mn, ops):
                for r in regs(tok) or []:
                    for v in st.get(r):
                        if not any(is_pointer(a) for a in v):
                            findings.append((ins[0], "%s %s" % (mn, " ".join(ops)), sorted(v)))
        if ("load" in mn or "atomic" in mn) and vreg_range(ops[0]) and len(address_operands(mn, ops)) and \
                address_operands(mn, ops)[0] != ops[0]:
            for k in vreg_range(ops[0]):
                st.r[k] = {MEM}
        return findings
    if mn.startswith(("buffer_load", "ds_read", "ds_load", "buffer_atomic", "ds_bpermute", "ds_permute", "ds_swizzle")):
        for k in vreg_range(ops[0]) or []:
            st.r[k] = {MEM}
        return findings
    if mn.startswith("scratch_"):
        off = next((int(o.split(":")[1], 0) for o in ops if o.startswith("offset:")), 0)
        if mn.startswith("scratch_store"):
            for k, r in enumerate(vreg_range(ops[1]) or []):
                st.r[("m", off + 4 * k)] = set(st.get(r))
        elif mn.startswith("scratch_load"):
            for k, r in enumerate(vreg_range(ops[0]) or []):
                st.r[r] = set(st.r.get(("m", off + 4 * k), {UNKNOWN}))
        return findings
    if mn.startswith(("s_load", "s_buffer_load")):
        dst, base = sreg_range(ops[0]), sreg_range(ops[1])
        off = imm(ops[2]) if len(ops) > 2 else None
        bv = set().union(*[st.get(r) for r in base]) if base else {UNKNOWN}
        for k, r in enumerate(dst or []):
            st.r[r] = {frozenset([load_atom(bv, off + 4 * k) if off is not None else ("other", -1)])}
        return findings
    if mn.startswith(("s_cmp", "s_bitcmp")):
        a, b = const_of(st.val(ops[0])), const_of(st.val(ops[1]))
        st.scc = eval_cmp(mn, a, b) if a is not None and b is not None and not mn.endswith("64") else None
        return findings
    if mn.startswith("s_"):
        return scalar(st, mn, ops) or findings
    if mn.startswith("v_"):
        vector(st, mn, ops, direction)
    return findings


def scalar(st, mn, ops):
    """transfer of a scalar ALU instruction"""
    if mn.startswith(NO_DEST) or "store" in mn or "atomic" in mn:  # no scalar register written
        return
    d, srcs = ops[0], ops[1:]
    keep_scc = mn.startswith(SCC_KEEP)
    if mn.endswith("saveexec_b64"):
        for r in sreg_range(d) or []:
            st.r[r] = {LANE}
        st.exec, st.scc = U, None
        return
    if d in ("exec", "vcc"):
        res = U
        if len(srcs) == 2 and mn in ("s_and_b64", "s_andn2_b64", "s_or_b64"):
            a, b = mask_state(st, srcs[0]), mask_state(st, srcs[1])
            if mn == "s_and_b64":
                res = mask_and(a, b)
            elif mn == "s_andn2_b64":
                res = mask_and(a, mask_not(b))
            else:
                res = NZ if {a, b} & {NZ, ONES} else (Z if a == b == Z else U)
            res = NZ if res == ONES else res
            st.scc = {Z: 0, NZ: 1}.get(res)
        elif mn in ("s_mov_b64", "s_cselect_b64") or not keep_scc:
            if mn == "s_mov_b64":
                res = mask_state(st, srcs[0])
                res = NZ if res == ONES else res
            if not keep_scc:
                st.scc = None
        setattr(st, "exec" if d == "exec" else "vcc", res)
        if d == "vcc":
            halves = operand_dwords(st, srcs[0], 2) if mn == "s_mov_b64" else [{LANE}] * 2
            st.r.update(zip(VCC_HALVES, halves))
        return
    if d.startswith("exec_"):
        st.exec = U
        if not keep_scc:
            st.scc = None
        return
    if d in VCC_HALVES:
        st.vcc = U
    dst = [d] if d in VCC_HALVES else sreg_range(d)
    if dst is None:
        if not keep_scc:
            st.scc = None
        return
    new = None
    if mn.startswith("s_cmov") and st.scc == 0:
        return
    if mn.startswith(("s_mov_b", "s_cmov_b", "s_movk")) or (mn.startswith("s_cselect") and st.scc is not None):
        src = srcs[0] if not mn.startswith("s_cselect") else srcs[0 if st.scc else 1]
        new = operand_dwords(st, src, len(dst))
        if mn.startswith("s_cmov") and st.scc is None:
            new = [n | st.get(r) for n, r in zip(new, dst)]
    elif mn.startswith("s_cselect"):
        new = [a | b for a, b in zip(operand_dwords(st, srcs[0], len(dst)), operand_dwords(st, srcs[1], len(dst)))]
    elif mn.startswith("s_getpc"):
        new = [{frozenset([("pc",)])}] * len(dst)
    elif len(dst) == 1 and len(srcs) == 2 and not mn.startswith(("s_addc", "s_subb")):
        a, b = const_of(st.val(srcs[0])), const_of(st.val(srcs[1]))
        r = eval_scalar(mn, a, b) if a is not None and b is not None else None
        if r is not None:
            new = [const(r[0])]
            if not keep_scc:
                st.scc, keep_scc = r[1], True
    if new is None:
        v = combine([st.val(s) for s in srcs if regs(s) or s in VCC_HALVES + ("vcc",)])
        new = [v] * len(dst)
    for r, v in zip(dst, new):
        st.r[r] = set(v)
    if not keep_scc:
        st.scc = None


def operand_dwords(st, tok, n):
    """per-dword provenance of a scalar source operand of n dwords"""
    rr = list(VCC_HALVES) if tok == "vcc" else sreg_range(tok)
    if rr and len(rr) == n:
        return [set(st.get(r)) for r in rr]
    c = imm(tok)
    if c is not None:
        return const64(c) if n == 2 else [const(c)]
    return [st.val(tok)] * n


def vcc_mask(st):
    """VCC written as a lane mask by a vector instruction"""
    st.vcc = U
    st.r.update((h, {LANE}) for h in VCC_HALVES)


def vector(st, mn, ops, direction):
    """transfer of a vector ALU instruction"""
    if mn.startswith("v_cmpx"):
        st.exec = U
        for r in sreg_range(ops[0]) or []:
            st.r[r] = {LANE}
        return
    if mn.startswith("v_cmp"):
        if ops[0] == "vcc" or mn.endswith("_e32"):
            vcc_mask(st)
        for r in sreg_range(ops[0]) or []:
            st.r[r] = {LANE}
        return
    srcs = ops[1:]
    if "_co_" in mn or mn.startswith(("v_addc", "v_subb", "v_div_scale", "v_mad_u64", "v_mad_i64")):
        if mn.endswith("_e32") and "_co_" in mn:
            vcc_mask(st)
        elif len(ops) > 1 and (ops[1] == "vcc" or sreg_range(ops[1])):
            if ops[1] == "vcc":
                vcc_mask(st)
            for r in sreg_range(ops[1]) or []:
                st.r[r] = {LANE}
            srcs = ops[2:]
        srcs = [o for o in srcs if o != "vcc"]
    d = ops[0]
    if mn.startswith(("v_readfirstlane", "v_readlane")):
        src = vreg_range(srcs[0])
        v = st.get(src[0]) if src else st.val(srcs[0])
        if mn.startswith("v_readlane") and src and len(srcs) > 1 and imm(srcs[1]) is not None:
            v = st.r.get(("vl", src[0][1], imm(srcs[1])), v)
        if v == {TIDY}:
            v = const(direction)
        for r in sreg_range(d) or []:
            st.r[r] = set(v)
        return
    if mn.startswith("v_writelane"):
        dst = vreg_range(d)
        if dst and len(srcs) > 1 and imm(srcs[1]) is not None:
            st.r[("vl", dst[0][1], imm(srcs[1]))] = set(st.val(srcs[0]))
        return
    dst = vreg_range(d)
    if not dst:
        for r in sreg_range(d) or []:  # VOP3 forms with a scalar destination
            st.r[r] = {LANE}
        if d == "vcc":
            vcc_mask(st)
        elif d in VCC_HALVES:
            st.vcc, st.r[d] = U, {LANE}
        return
    if mn.startswith(("v_mov_b32", "v_mov_b64")):
        rr = regs(srcs[0])
        if rr and len(rr) == len(dst):
            vals = [set(st.get(r)) for r in rr]
        elif imm(srcs[0]) is not None and len(dst) == 2:
            vals = const64(imm(srcs[0]))
        else:
            vals = [st.val(srcs[0])] * len(dst)
    elif mn.startswith("v_cndmask"):
        vals = [st.val(srcs[0]) | st.val(srcs[1])] * len(dst)
    elif mn.startswith("v_bfe_u32") and st.val(srcs[0]) == {TID} and imm(srcs[1]) == 10 and imm(srcs[2]) == 10:
        vals = [{TIDY}] * len(dst)
    else:
        sv = [st.val(s) for s in srcs if regs(s) or s in VCC_HALVES + ("vcc",)]
        plain = all(a[0] in PLAIN_ATOMS for s in sv for v in s for a in v)
        vals = [{LANE} if plain else combine(sv)] * len(dst)
    for r, v in zip(dst, vals):
        st.r[r] = v


def entry_state(setup, direction):
    """register provenances at kernel entry, from the descriptor's SGPR set-up"""
    user, tgx, tgy, kernarg_at = setup
    r = {("v", 0): {TID}}
    if kernarg_at is not None:
        r[kernarg_at] = r[kernarg_at + 1] = {frozenset([("kernarg",)])}
    if tgx:
        r[user] = {frozenset([("wgid",)])}
    if tgy:
        r[user + int(tgx)] = const(direction)
    return State(r)


def successors(st, ins, fall, tb):
    """(block, state) pairs after a block's last instruction, deciding the branches that the state decides"""
    _, mn, _, _ = ins
    if mn.startswith(("s_endpgm", "s_setpc")):
        return []
    if mn.startswith("s_branch"):
        return [(tb, st)]
    if not mn.startswith("s_cbranch"):
        return [(fall, st)]
    cond = mn[len("s_cbranch_"):]
    if cond in ("scc0", "scc1"):
        want = int(cond == "scc1")
        if st.scc is not None:
            return [(tb if st.scc == want else fall, st)]
        t, f = st.copy(), st.copy()
        t.scc, f.scc = want, 1 - want
        return [(tb, t), (fall, f)]
    if cond in ("vccz", "vccnz", "execz", "execnz"):
        reg = cond[:-1].rstrip("n") if cond.endswith("z") else cond
        reg = "vcc" if cond.startswith("vcc") else "exec"
        on_zero = not cond.endswith("nz")
        cur = getattr(st, reg)
        if cur != U:
            return [(tb if (cur == Z) == on_zero else fall, st)]
        t, f = st.copy(), st.copy()
        setattr(t, reg, Z if on_zero else NZ)
        setattr(f, reg, NZ if on_zero else Z)
        return [(tb, t), (fall, f)]
    return [(tb, st.copy()), (fall, st)]


def check_kernel(instrs, setup, fam=STAGE_FAMILY):
    """(sorted findings over every direction: (address, text, provenance, direction),
    number of global accesses checked on some path with active lanes); fam =
    (directions, pointer argument offsets), the stage kernels' by default"""
    global pointer_dwords
    directions, ptrs = fam
    pointer_dwords = frozenset(o + k for o in ptrs for k in (0, 4))
    blocks, bid = blocks_of(instrs)
    found, checked = set(), set()
    for direction in range(directions):
        ins_state = [dict() for _ in blocks]  # block -> {EXEC state: State}
        st0 = entry_state(setup, direction)
        ins_state[0][st0.exec] = st0
        work = [(0, st0.exec)]
        while work:
            b, key = work.pop()
            st = ins_state[b][key].copy()
            s, e = blocks[b]
            for i in range(s, e):
                transfer(st, instrs[i], direction)
            fall = b + 1 if e < len(instrs) else None
            tgt = instrs[e - 1][3]
            for nb, nst in successors(st, instrs[e - 1], fall, bid.get(tgt) if tgt is not None else None):
                if nb is None:
                    continue
                cur = ins_state[nb].get(nst.exec)
                if cur is None:
                    ins_state[nb][nst.exec] = nst.copy()
                    work.append((nb, nst.exec))
                elif cur.join(nst):
                    work.append((nb, nst.exec))
        for b, (s, e) in enumerate(blocks):
            for st in ins_state[b].values():
                st = st.copy()
                for i in range(s, e):
                    if instrs[i][1].startswith(("global_", "flat_")) and st.exec != Z:
                        checked.add(i)
                    for addr, text, prov in transfer(st, instrs[i], direction):
                        found.add((addr, text, tuple(prov), direction))
    return sorted(found), len(checked)


def kernel_setup(elf):
    """{kernel: (user SGPR count, workgroup id x enabled, workgroup id y enabled, kernarg pointer SGPR or None)}
    from the 64-byte kernel descriptors (<name>.kd) of an AMDGPU code object"""
    with open(elf, "rb") as f:
        data = f.read()
    shoff, = struct.unpack_from("<Q", data, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", data, 0x3A)
    secs = [struct.unpack_from("<IIQQQQIIQQ", data, shoff + k * shentsize) for k in range(shnum)]
    symtab = next(s for s in secs if s[1] == 2)  # SHT_SYMTAB
    strtab = secs[symtab[6]]
    out = {}
    for off in range(symtab[4], symtab[4] + symtab[5], 24):
        name_off, _, _, shndx, value, _ = struct.unpack_from("<IBBHQQ", data, off)
        start = strtab[4] + name_off
        name = data[start:data.index(b"\0", start)].decode()
        if not name.endswith(".kd") or shndx >= len(secs):
            continue
        sec = secs[shndx]
        kd = sec[4] + value - sec[3]
        rsrc2, = struct.unpack_from("<I", data, kd + 0x34)
        props, = struct.unpack_from("<H", data, kd + 0x38)
        user = (rsrc2 >> 1) & 31
        # user SGPRs in order: private segment buffer (4), dispatch ptr (2), queue ptr (2), kernarg ptr (2)
        k = 4 * (props & 1) + 2 * ((props >> 1) & 1) + 2 * ((props >> 2) & 1)
        out[name[:-3]] = (user, bool(rsrc2 >> 7 & 1), bool(rsrc2 >> 8 & 1), k if props >> 3 & 1 else None)
    return out


def disassemble(obj):
    """(disassembly text, kernel set-up) of the gfx950 code object embedded in a host object"""
    with tempfile.TemporaryDirectory() as d:
        fb, elf = os.path.join(d, "fatbin"), os.path.join(d, "gfx950.elf")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fb, obj, os.devnull],
                       check=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", "--input=" + fb,
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + elf], check=True)
        text = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--mcpu=gfx950", elf], check=True,
                              capture_output=True, text=True).stdout
        return text, kernel_setup(elf)


def check_object(obj, workers=8):
    """{kernel: (findings, accesses checked)} for every checked kernel in a host object"""
    text, setup = disassemble(obj)
    kernels = parse(text)
    names = sorted(kernels)
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers) as pool:
        results = pool.map(check_kernel, [kernels[n] for n in names], [setup[n] for n in names],
                           [family(n) for n in names])
        return dict(zip(names, results))


def main():
    obj = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "_build", "kernels.hip.o")
    results = check_object(obj)
    bad = checked = 0
    for name, (findings, n) in results.items():
        checked += n
        for addr, text, prov, direction in findings:
            bad += 1
            print("%s direction %d @%x: %s: address may hold %s" % (name, direction, addr, text, list(prov)))
    print("check_load_bases: %d kernels (RK3 stage, line products, clover), %d global accesses checked, "
          "%d through a non-pointer address" % (len(results), checked, bad))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
