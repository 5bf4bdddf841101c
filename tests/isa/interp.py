"""An interpreter for the gfx950 assembly of one compiled kernel (the instructions the wave-streamed SOR
kernel uses): scalar and vector ALU, DPP wave_shr / row_shl moves, lane read/write, global / buffer / LDS
memory, exec-mask control flow and long branches.

It runs several waves of the compiled code on the CPU, interleaved at random points (every s_sleep
yields), over a flat memory arena that faults on out-of-bounds addresses, and reports runs that exceed an
instruction budget.  Memory is sequentially consistent, so it checks what the compiler made of the kernel's
control and data flow (a loop that drops lanes, a wrong ring index, an address outside its buffer), not
the chip's memory ordering.
"""
import re
import struct
import sys

import numpy as np

MASK32 = 0xFFFFFFFF
MASK64 = 0xFFFFFFFFFFFFFFFF
VCC, EXEC = 106, 126


def parse_kernel(path, name):
    lines = open(path).read().split("\n")
    start = None
    for i, l in enumerate(lines):
        if l.startswith(name + ":"):
            start = i
            break
    assert start is not None, name
    prog, labels = [], {}
    for l in lines[start + 1:]:
        s = l.split(";")[0].strip()
        if not s:
            continue
        if s.startswith(".Lfunc_end") or s.startswith(".section") or s == ".end_amdhsa_kernel":
            break
        if s.endswith(":"):
            labels[s[:-1]] = len(prog)
            continue
        if s.startswith("."):
            continue
        parts = s.split(None, 1)
        op = parts[0]
        ops = []
        if len(parts) > 1:
            ops = [o.strip() for o in parts[1].split(",")]
        # split modifiers that follow the last operand separated by spaces
        mods = {}
        if ops:
            last = ops[-1].split()
            ops[-1] = last[0]
            for m in last[1:]:
                if ":" in m:
                    k, v = m.split(":", 1)
                    mods[k] = v
                else:
                    mods[m] = True
        if op == "s_setpc_b64":
            # long branch: s_getpc_b64 / s_add_u32 (target-post)&... / s_addc_u32 / s_setpc_b64
            add = prog[-2]
            m = re.search(r"\((\.LBB[0-9_]+)-", add[1][2])
            assert add[0] == "s_add_u32" and m, add
            prog[-3] = ("s_nop", [], {}, "long branch")
            prog[-2] = ("s_nop", [], {}, "long branch")
            prog[-1] = ("s_nop", [], {}, "long branch")
            prog.append(("s_branch", [m.group(1)], {}, "long branch to " + m.group(1)))
            continue
        prog.append((op, ops, mods, s))
    return prog, labels


FLOAT_CONST = {"0.5": 0.5, "-0.5": -0.5, "1.0": 1.0, "-1.0": -1.0, "2.0": 2.0, "-2.0": -2.0, "4.0": 4.0,
               "-4.0": -4.0}


def reg(o):
    m = re.fullmatch(r"([sva])(\d+)", o)
    if m:
        return m.group(1), int(m.group(2)), 1
    m = re.fullmatch(r"([sva])\[(\d+):(\d+)\]", o)
    if m:
        return m.group(1), int(m.group(2)), int(m.group(3)) - int(m.group(2)) + 1
    if o == "vcc":
        return "s", VCC, 2
    if o == "exec":
        return "s", EXEC, 2
    if o == "vcc_lo":
        return "s", VCC, 1
    if o == "vcc_hi":
        return "s", VCC + 1, 1
    return None


class Fault(Exception):
    pass


class Machine:
    def __init__(self, arena_bytes):
        self.mem = np.zeros(arena_bytes, np.uint8)
        self.base = 0x100000
        self.top = self.base
        self.clock = 0

    def alloc(self, nbytes, align=256):
        self.top = (self.top + align - 1) // align * align
        a = self.top
        self.top += nbytes
        assert self.top - self.base <= self.mem.size
        return a

    def off(self, addr, n):
        o = addr - self.base
        if o < 0 or o + n > self.top - self.base:
            raise Fault("access out of bounds: addr 0x%x n %d" % (addr, n))
        return o

    def rd32(self, addr):
        o = self.off(addr, 4)
        return int(self.mem[o:o + 4].view(np.uint32)[0])

    def wr32(self, addr, v):
        o = self.off(addr, 4)
        self.mem[o:o + 4].view(np.uint32)[0] = v & MASK32


class Wave:
    def __init__(self, m, prog, labels, kernarg, wgid, nvgpr=512):
        self.m, self.prog, self.labels = m, prog, labels
        self.s = [0] * 128
        self.scc = 0
        self.v = np.zeros((nvgpr, 64), np.uint32)
        self.a = np.zeros((256, 64), np.uint32)
        self.v[0] = np.arange(64, dtype=np.uint32)
        self.s[0] = kernarg & MASK32
        self.s[1] = kernarg >> 32
        self.s[2] = wgid
        self.set_s64(EXEC, MASK64)
        self.pc = 0
        self.done = False
        self.count = 0
        self.lds = np.zeros(65536, np.uint8)

    # ---------------- register helpers
    def s64(self, i):
        return self.s[i] | (self.s[i + 1] << 32)

    def set_s64(self, i, v):
        v &= MASK64
        self.s[i] = v & MASK32
        self.s[i + 1] = v >> 32

    def execmask(self):
        e = self.s64(EXEC)
        return ((e >> np.arange(64, dtype=np.uint64)) & 1).astype(bool)

    @staticmethod
    def tomask(bits):
        return int(np.sum(bits.astype(np.uint64) << np.arange(64, dtype=np.uint64)))

    @staticmethod
    def frommask(v):
        return ((np.uint64(v) >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)

    def lit(self, o, width, fp=False):
        if o in FLOAT_CONST:
            f = FLOAT_CONST[o]
            if width == 2:
                return struct.unpack("<Q", struct.pack("<d", f))[0] if fp else struct.unpack("<I", struct.pack("<f", f))[0]
            return struct.unpack("<I", struct.pack("<f", f))[0]
        v = int(o, 0)
        if width == 2:
            if v < 0 or (fp is False and v <= 0xFFFFFFFF and o.startswith("0x") and False):
                return v & MASK64
            if fp and -16 <= v <= 64:
                return struct.unpack("<Q", struct.pack("<d", float(v)))[0]
            return v & MASK64
        return v & MASK32

    def sread(self, o, width=1, fp=False):
        r = reg(o)
        if r:
            k, i, n = r
            assert k == "s", o
            if width == 2:
                return self.s64(i)
            return self.s[i]
        if o == "scc":
            return self.scc
        return self.lit(o, width, fp)

    def swrite(self, o, val, width=1):
        k, i, n = reg(o)
        assert k == "s"
        if width == 2:
            self.set_s64(i, val)
        else:
            self.s[i] = val & MASK32

    def vread(self, o, width=1, fp=False):
        """Per-lane values: uint32 array (width 1) or uint64 array (width 2).  Float negation / abs modifiers."""
        if o.startswith("-") and (reg(o[1:]) or o[1:].startswith("|")):
            x = self.vread(o[1:], width, fp)
            return x ^ (np.uint64(1 << 63) if width == 2 else np.uint32(1 << 31))
        if o.startswith("|") and o.endswith("|"):
            x = self.vread(o[1:-1], width, fp)
            return x & (np.uint64(MASK64 >> 1) if width == 2 else np.uint32(MASK32 >> 1))
        r = reg(o)
        if r:
            k, i, n = r
            if k in "va":
                F = self.v if k == "v" else self.a
                if width == 2:
                    return F[i].astype(np.uint64) | (F[i + 1].astype(np.uint64) << np.uint64(32))
                return F[i].copy()
            if width == 2:
                return np.full(64, self.s64(i), np.uint64)
            return np.full(64, self.s[i], np.uint32)
        val = self.lit(o, width, fp)
        return np.full(64, val, np.uint64 if width == 2 else np.uint32)

    def vwrite(self, o, vals, width=1, mask=None):
        k, i, n = reg(o)
        assert k in "va"
        F = self.v if k == "v" else self.a
        e = self.execmask() if mask is None else mask
        if width == 2:
            vals = vals.astype(np.uint64)
            F[i][e] = (vals & np.uint64(MASK32)).astype(np.uint32)[e]
            F[i + 1][e] = (vals >> np.uint64(32)).astype(np.uint32)[e]
        else:
            F[i][e] = vals.astype(np.uint32)[e]

    def vfile(self, o):
        k, i, n = reg(o)
        return (self.v if k == "v" else self.a), i

    # ---------------- memory
    def gaddr(self, ops, mods):
        """global_* addressing: (vaddr, saddr|off) -> per-lane 64-bit addresses"""
        vaddr, saddr = ops
        off = int(mods.get("offset", "0"), 0)
        if saddr == "off":
            a = self.vread(vaddr, 2)
        else:
            a = self.vread(vaddr, 1).astype(np.uint64) + np.uint64(self.s64(reg(saddr)[1]))
        return a + np.uint64(off & MASK64)

    # ---------------- execution
    def step(self):
        op, ops, mods, text = self.prog[self.pc]
        self.pc += 1
        self.count += 1
        f = getattr(self, "op_" + op, None)
        if f is None:
            raise NotImplementedError(text)
        return f(ops, mods)

    # scalar ALU
    def op_s_nop(self, o, m): pass
    def op_s_waitcnt(self, o, m): pass
    def op_buffer_wbl2(self, o, m): pass
    def op_buffer_inv(self, o, m): pass
    def op_s_barrier(self, o, m): pass

    def op_s_endpgm(self, o, m):
        self.done = True

    def op_s_sleep(self, o, m):
        return "yield"

    def op_s_memrealtime(self, o, m):
        self.swrite(o[0], self.m.clock, 2)

    def op_s_load_dword(self, o, m): self._sload(o, 1)
    def op_s_load_dwordx8(self, o, m): self._sload(o, 8)
    def op_s_load_dwordx2(self, o, m): self._sload(o, 2)
    def op_s_load_dwordx4(self, o, m): self._sload(o, 4)
    def op_s_load_dwordx16(self, o, m): self._sload(o, 16)

    def _sload(self, o, n):
        k, i, nn = reg(o[0])
        base = self.s64(reg(o[1])[1]) + int(o[2], 0)
        for j in range(n):
            self.s[i + j] = self.m.rd32(base + 4 * j)

    def _s32(self, o, fn, scc=None):
        a, b = self.sread(o[1]), self.sread(o[2])
        r = fn(a, b)
        self.swrite(o[0], r & MASK32)
        if scc is not None:
            self.scc = int(scc(a, b, r))

    @staticmethod
    def sx(v):
        v &= MASK32
        return v - (1 << 32) if v & 0x80000000 else v

    def op_s_add_i32(self, o, m): self._s32(o, lambda a, b: a + b)
    def op_s_sub_i32(self, o, m): self._s32(o, lambda a, b: a - b)
    def op_s_mul_i32(self, o, m): self._s32(o, lambda a, b: self.sx(a) * self.sx(b))
    def op_s_mul_hi_i32(self, o, m): self._s32(o, lambda a, b: (self.sx(a) * self.sx(b)) >> 32)
    def op_s_mul_hi_u32(self, o, m): self._s32(o, lambda a, b: (a * b) >> 32)
    def op_s_add_u32(self, o, m): self._s32(o, lambda a, b: a + b, lambda a, b, r: r > MASK32)
    def op_s_sub_u32(self, o, m): self._s32(o, lambda a, b: a - b, lambda a, b, r: b > a)

    def op_s_addc_u32(self, o, m):
        c = self.scc
        self._s32(o, lambda a, b: a + b + c, lambda a, b, r: r > MASK32)

    def op_s_subb_u32(self, o, m):
        c = self.scc
        self._s32(o, lambda a, b: a - b - c, lambda a, b, r: b + c > a)

    def op_s_ashr_i32(self, o, m): self._s32(o, lambda a, b: self.sx(a) >> (b & 31), lambda a, b, r: (r & MASK32) != 0)
    def op_s_lshr_b32(self, o, m): self._s32(o, lambda a, b: a >> (b & 31), lambda a, b, r: (r & MASK32) != 0)
    def op_s_lshl_b32(self, o, m): self._s32(o, lambda a, b: a << (b & 31), lambda a, b, r: (r & MASK32) != 0)
    def op_s_or_b32(self, o, m): self._s32(o, lambda a, b: a | b, lambda a, b, r: r != 0)
    def op_s_xor_b32(self, o, m): self._s32(o, lambda a, b: a ^ b, lambda a, b, r: r != 0)
    def op_s_and_b32(self, o, m): self._s32(o, lambda a, b: a & b, lambda a, b, r: r != 0)
    def op_s_min_i32(self, o, m): self._s32(o, lambda a, b: a if self.sx(a) < self.sx(b) else b, lambda a, b, r: self.sx(a) < self.sx(b))
    def op_s_max_i32(self, o, m): self._s32(o, lambda a, b: a if self.sx(a) > self.sx(b) else b, lambda a, b, r: self.sx(a) > self.sx(b))

    def op_s_abs_i32(self, o, m):
        a = self.sx(self.sread(o[1]))
        r = abs(a) & MASK32
        self.swrite(o[0], r)
        self.scc = int(r != 0)

    def op_s_mov_b32(self, o, m): self.swrite(o[0], self.sread(o[1]))
    def op_s_mov_b64(self, o, m): self.swrite(o[0], self.sread(o[1], 2), 2)

    def _s64(self, o, fn):
        a, b = self.sread(o[1], 2), self.sread(o[2], 2)
        r = fn(a, b) & MASK64
        self.swrite(o[0], r, 2)
        self.scc = int(r != 0)

    def op_s_and_b64(self, o, m): self._s64(o, lambda a, b: a & b)
    def op_s_or_b64(self, o, m): self._s64(o, lambda a, b: a | b)
    def op_s_xor_b64(self, o, m): self._s64(o, lambda a, b: a ^ b)
    def op_s_andn2_b64(self, o, m): self._s64(o, lambda a, b: a & ~b)

    def op_s_lshl_b64(self, o, m):
        a, b = self.sread(o[1], 2), self.sread(o[2])
        r = (a << (b & 63)) & MASK64
        self.swrite(o[0], r, 2)
        self.scc = int(r != 0)

    def op_s_and_saveexec_b64(self, o, m):
        e = self.s64(EXEC)
        src = self.sread(o[1], 2)
        self.swrite(o[0], e, 2)
        ne = src & e
        self.set_s64(EXEC, ne)
        self.scc = int(ne != 0)

    def op_s_bcnt1_i32_b64(self, o, m):
        r = bin(self.sread(o[1], 2)).count("1")
        self.swrite(o[0], r)
        self.scc = int(r != 0)

    def _scmp(self, o, fn, signed):
        a, b = self.sread(o[0]), self.sread(o[1])
        if signed:
            a, b = self.sx(a), self.sx(b)
        self.scc = int(fn(a, b))

    def op_s_cmp_ge_i32(self, o, m): self._scmp(o, lambda a, b: a >= b, True)
    def op_s_cmp_gt_i32(self, o, m): self._scmp(o, lambda a, b: a > b, True)
    def op_s_cmp_lt_i32(self, o, m): self._scmp(o, lambda a, b: a < b, True)
    def op_s_cmp_ge_u32(self, o, m): self._scmp(o, lambda a, b: a >= b, False)
    def op_s_cmp_eq_u32(self, o, m): self._scmp(o, lambda a, b: a == b, False)

    def op_s_cselect_b32(self, o, m): self.swrite(o[0], self.sread(o[1]) if self.scc else self.sread(o[2]))
    def op_s_cselect_b64(self, o, m): self.swrite(o[0], self.sread(o[1], 2) if self.scc else self.sread(o[2], 2), 2)

    def _br(self, o, cond):
        if cond:
            self.pc = self.labels[o[0]]

    def op_s_branch(self, o, m): self._br(o, True)
    def op_s_cbranch_scc0(self, o, m): self._br(o, self.scc == 0)
    def op_s_cbranch_scc1(self, o, m): self._br(o, self.scc == 1)
    def op_s_cbranch_vccz(self, o, m): self._br(o, self.s64(VCC) == 0)
    def op_s_cbranch_vccnz(self, o, m): self._br(o, self.s64(VCC) != 0)
    def op_s_cbranch_execz(self, o, m): self._br(o, self.s64(EXEC) == 0)
    def op_s_cbranch_execnz(self, o, m): self._br(o, self.s64(EXEC) != 0)

    # vector ALU
    def op_v_mov_b32_e32(self, o, m): self.vwrite(o[0], self.vread(o[1]))
    def op_v_mov_b64_e32(self, o, m): self.vwrite(o[0], self.vread(o[1], 2, fp=True), 2)

    def op_v_mov_b32_dpp(self, o, m):
        assert m.get("row_mask") == "0xf" and m.get("bank_mask") == "0xf", m
        assert "bound_ctrl" not in m, m
        e = self.execmask()
        src = self.vread(o[1])
        F, i = self.vfile(o[0])
        new = F[i].copy()
        lanes = np.arange(64)
        if m.get("wave_shr") == "1":
            srcl = lanes - 1
            valid = srcl >= 0
        elif "row_shl" in m:
            j = int(m["row_shl"])
            srcl = lanes + j
            valid = (srcl // 16) == (lanes // 16)
        else:
            raise NotImplementedError(str(m))
        srcl = np.clip(srcl, 0, 63)
        valid = valid & e[srcl]
        w = e & valid
        new[w] = src[srcl][w]
        F[i] = new

    def op_v_readlane_b32(self, o, m):
        k, i, n = reg(o[1])
        self.swrite(o[0], int(self.v[i][int(self.sread(o[2]))]))

    def op_v_writelane_b32(self, o, m):
        k, i, n = reg(o[0])
        self.v[i][int(self.sread(o[2]))] = self.sread(o[1])

    def op_v_readfirstlane_b32(self, o, m):
        e = self.execmask()
        k, i, n = reg(o[1])
        lane = int(np.argmax(e)) if e.any() else 0
        self.swrite(o[0], int(self.v[i][lane]))

    def op_v_mbcnt_lo_u32_b32(self, o, m):
        msk = self.sread(o[1]) & MASK32
        lanes = np.arange(64)
        cnt = np.array([bin(msk & ((1 << min(l, 32)) - 1)).count("1") for l in lanes], np.uint32)
        self.vwrite(o[0], cnt + self.vread(o[2]))

    def op_v_mbcnt_hi_u32_b32(self, o, m):
        msk = self.sread(o[1]) & MASK32
        lanes = np.arange(64)
        cnt = np.array([bin(msk & ((1 << max(l - 32, 0)) - 1)).count("1") for l in lanes], np.uint32)
        self.vwrite(o[0], cnt + self.vread(o[2]))

    def _vcmp(self, o, fn, width=1, signed=False, e64=False):
        a, b = self.vread(o[1], width), self.vread(o[2], width)
        if signed:
            a, b = a.astype(np.int32 if width == 1 else np.int64), b.astype(np.int32 if width == 1 else np.int64)
        r = fn(a, b) & self.execmask()
        self.swrite(o[0], self.tomask(r), 2)

    def __getattr__(self, name):
        mm = re.fullmatch(r"op_v_cmp_(eq|ne|lt|le|gt|ge)_([iu])(32|64)_e(32|64)", name)
        if mm:
            rel, sg, w, _ = mm.groups()
            fn = {"eq": np.equal, "ne": np.not_equal, "lt": np.less, "le": np.less_equal, "gt": np.greater,
                  "ge": np.greater_equal}[rel]
            return lambda o, m: self._vcmp(o, fn, 2 if w == "64" else 1, sg == "i")
        raise AttributeError(name)

    def _vmask(self, o):
        return self.frommask(self.sread(o, 2))

    def op_v_cndmask_b32_e32(self, o, m):
        c = self._vmask(o[3])
        self.vwrite(o[0], np.where(c, self.vread(o[2]), self.vread(o[1])))

    op_v_cndmask_b32_e64 = op_v_cndmask_b32_e32

    def op_v_add_u32_e32(self, o, m): self.vwrite(o[0], (self.vread(o[1]).astype(np.uint64) + self.vread(o[2])) & MASK32)
    def op_v_sub_u32_e32(self, o, m): self.vwrite(o[0], (self.vread(o[1]).astype(np.int64) - self.vread(o[2])) & MASK32)

    def op_v_subrev_co_u32_e64(self, o, m):
        a, b = self.vread(o[2]).astype(np.int64), self.vread(o[3]).astype(np.int64)
        r = b - a
        self.vwrite(o[0], (r & MASK32).astype(np.uint64))
        self.swrite(o[1], self.tomask((r < 0) & self.execmask()), 2)

    def op_v_lshlrev_b32_e32(self, o, m):
        sh = self.vread(o[1]) & 31
        self.vwrite(o[0], (self.vread(o[2]).astype(np.uint64) << sh.astype(np.uint64)) & MASK32)

    def op_v_ashrrev_i32_e32(self, o, m):
        sh = (self.vread(o[1]) & 31).astype(np.int32)
        self.vwrite(o[0], (self.vread(o[2]).astype(np.int32) >> sh).astype(np.uint32))

    def op_v_lshlrev_b64(self, o, m):
        sh = (self.vread(o[1]) & 63).astype(np.uint64)
        self.vwrite(o[0], self.vread(o[2], 2) << sh, 2)

    def op_v_not_b32_e32(self, o, m): self.vwrite(o[0], ~self.vread(o[1]))
    def op_v_or_b32_e32(self, o, m): self.vwrite(o[0], self.vread(o[1]) | self.vread(o[2]))

    def op_v_alignbit_b32(self, o, m):
        a, b, c = self.vread(o[1]).astype(np.uint64), self.vread(o[2]).astype(np.uint64), self.vread(o[3]) & 31
        self.vwrite(o[0], (((a << np.uint64(32)) | b) >> c.astype(np.uint64)) & np.uint64(MASK32))

    def op_v_lshl_add_u64(self, o, m):
        a = self.vread(o[1], 2)
        sh = (self.vread(o[2]) & 7).astype(np.uint64)
        c = self.vread(o[3], 2)
        self.vwrite(o[0], (a << sh) + c, 2)

    def op_v_mad_u64_u32(self, o, m):
        a, b = self.vread(o[2]).astype(np.uint64), self.vread(o[3]).astype(np.uint64)
        c = self.vread(o[4], 2)
        self.vwrite(o[0], a * b + c, 2)
        self.swrite(o[1], 0, 2)

    def op_v_mad_i64_i32(self, o, m):
        a, b = self.vread(o[2]).astype(np.int32).astype(np.int64), self.vread(o[3]).astype(np.int32).astype(np.int64)
        c = self.vread(o[4], 2).astype(np.int64)
        self.vwrite(o[0], (a * b + c).astype(np.uint64), 2)
        self.swrite(o[1], 0, 2)

    @staticmethod
    def _imm16(x):
        v = int(x, 0)
        return v - 0x10000 if v >= 0x8000 else v

    def op_s_addk_i32(self, o, m): self.swrite(o[0], (self.sx(self.sread(o[0])) + self._imm16(o[1])) & MASK32)
    def op_s_movk_i32(self, o, m): self.swrite(o[0], self._imm16(o[1]) & MASK32)
    def op_v_and_b32_e32(self, o, m): self.vwrite(o[0], self.vread(o[1]) & self.vread(o[2]))
    def op_v_lshrrev_b32_e32(self, o, m): self.vwrite(o[0], self.vread(o[2]) >> (self.vread(o[1]) & 31))
    def op_v_lshl_or_b32(self, o, m):
        self.vwrite(o[0], ((self.vread(o[1]).astype(np.uint64) << (self.vread(o[2]) & 31).astype(np.uint64)) & MASK32) | self.vread(o[3]))
    def op_v_mad_u32_u24(self, o, m):
        a, b = self.vread(o[1]) & 0xFFFFFF, self.vread(o[2]) & 0xFFFFFF
        self.vwrite(o[0], (a.astype(np.uint64) * b + self.vread(o[3])) & MASK32)
    def op_v_mul_f64(self, o, m):
        self.vwrite(o[0], (self.f64(self.vread(o[1], 2, True)) * self.f64(self.vread(o[2], 2, True))).view(np.uint64), 2)

    def op_s_mulk_i32(self, o, m):
        self.swrite(o[0], (self.sx(self.sread(o[0])) * self.sx(self.lit(o[1], 1) if not o[1].startswith("0x") else int(o[1], 16) if int(o[1], 16) < 0x8000 else int(o[1], 16) - 0x10000)) & MASK32)

    def op_v_accvgpr_write_b32(self, o, m): self.vwrite(o[0], self.vread(o[1]))
    def op_v_accvgpr_read_b32(self, o, m): self.vwrite(o[0], self.vread(o[1]))
    def op_v_accvgpr_mov_b32(self, o, m): self.vwrite(o[0], self.vread(o[1]))
    def op_v_subrev_u32_e32(self, o, m): self.vwrite(o[0], (self.vread(o[2]).astype(np.int64) - self.vread(o[1])) & MASK32)

    def op_v_add_f32_e32(self, o, m):
        self.vwrite(o[0], (self.f32(self.vread(o[1])) + self.f32(self.vread(o[2]))).view(np.uint32))

    def op_v_sub_f32_e32(self, o, m):
        self.vwrite(o[0], (self.f32(self.vread(o[1])) - self.f32(self.vread(o[2]))).view(np.uint32))

    def op_v_fma_f32(self, o, m):
        a, b, c = (self.f32(self.vread(x)).astype(np.float64) for x in o[1:4])
        self.vwrite(o[0], (a * b + c).astype(np.float32).view(np.uint32))

    def op_v_fmac_f32_e32(self, o, m):
        a, b = self.f32(self.vread(o[1])).astype(np.float64), self.f32(self.vread(o[2])).astype(np.float64)
        c = self.f32(self.vread(o[0])).astype(np.float64)
        self.vwrite(o[0], (a * b + c).astype(np.float32).view(np.uint32))

    def op_v_cvt_f32_f64_e32(self, o, m):
        self.vwrite(o[0], self.f64(self.vread(o[1], 2, True)).astype(np.float32).view(np.uint32))

    def op_v_cvt_f64_f32_e32(self, o, m):
        self.vwrite(o[0], self.f32(self.vread(o[1])).astype(np.float64).view(np.uint64), 2)

    def op_v_min_i32_e32(self, o, m):
        self.vwrite(o[0], np.minimum(self.vread(o[1]).astype(np.int32), self.vread(o[2]).astype(np.int32)).astype(np.uint32))

    # float
    @staticmethod
    def f64(x): return x.view(np.float64)
    @staticmethod
    def f32(x): return x.view(np.float32)

    def op_v_add_f64(self, o, m):
        self.vwrite(o[0], (self.f64(self.vread(o[1], 2, True)) + self.f64(self.vread(o[2], 2, True))).view(np.uint64), 2)

    def op_v_fma_f64(self, o, m):
        a, b, c = (self.f64(self.vread(x, 2, True)) for x in o[1:4])
        self.vwrite(o[0], (a.astype(np.longdouble) * b + c).astype(np.float64).view(np.uint64), 2)

    def op_v_fmac_f64_e32(self, o, m):
        a, b = self.f64(self.vread(o[1], 2, True)), self.f64(self.vread(o[2], 2, True))
        c = self.f64(self.vread(o[0], 2, True))
        self.vwrite(o[0], (a.astype(np.longdouble) * b + c).astype(np.float64).view(np.uint64), 2)

    def op_v_mul_f32_e32(self, o, m):
        self.vwrite(o[0], (self.f32(self.vread(o[1])) * self.f32(self.vread(o[2]))).view(np.uint32))

    def op_v_cvt_f32_u32_e32(self, o, m): self.vwrite(o[0], self.vread(o[1]).astype(np.float32).view(np.uint32))

    def op_v_cvt_u32_f32_e32(self, o, m):
        f = self.f32(self.vread(o[1]))
        self.vwrite(o[0], np.clip(np.nan_to_num(f), 0, 4294967295.0).astype(np.uint64).astype(np.uint32))

    def op_v_rcp_iflag_f32_e32(self, o, m):
        with np.errstate(divide="ignore"):
            self.vwrite(o[0], (np.float32(1.0) / self.f32(self.vread(o[1]))).view(np.uint32))

    # memory
    def _gload(self, o, m, n):
        addr = self.gaddr(o[1:3], m)
        e = self.execmask()
        F, i = self.vfile(o[0])
        for l in np.nonzero(e)[0]:
            for j in range(n):
                F[i + j][l] = self.m.rd32(int(addr[l]) + 4 * j)

    def op_global_load_dword(self, o, m): self._gload(o, m, 1)
    def op_global_load_dwordx2(self, o, m): self._gload(o, m, 2)
    def op_global_load_dwordx4(self, o, m): self._gload(o, m, 4)

    def _gstore(self, o, m, n):
        addr = self.gaddr([o[0], o[2]], m)
        e = self.execmask()
        F, i = self.vfile(o[1])
        for l in np.nonzero(e)[0]:
            for j in range(n):
                self.m.wr32(int(addr[l]) + 4 * j, int(F[i + j][l]))

    def op_global_store_dword(self, o, m): self._gstore(o, m, 1)
    def op_global_store_dwordx2(self, o, m): self._gstore(o, m, 2)
    def op_global_store_dwordx4(self, o, m): self._gstore(o, m, 4)

    def _lds_addr(self, o, m):
        return self.vread(o) + np.uint32(int(m.get("offset", "0"), 0))

    def op_ds_write_b32(self, o, m):
        a = self._lds_addr(o[0], m)
        d = self.vread(o[1])
        for l in np.nonzero(self.execmask())[0]:
            self.lds[int(a[l]):int(a[l]) + 4].view(np.uint32)[0] = d[l]

    def op_ds_read_b32(self, o, m):
        a = self._lds_addr(o[1], m)
        k, i, n = reg(o[0])
        for l in np.nonzero(self.execmask())[0]:
            self.v[i][l] = self.lds[int(a[l]):int(a[l]) + 4].view(np.uint32)[0]

    def _buf(self, o, m):
        vdata, voff, rs, soff = o
        k, i, n = reg(rs)
        base = self.s[i] | ((self.s[i + 1] & 0xFFFF) << 32)
        nrec = self.s[i + 2]
        off = self.vread(voff).astype(np.uint64) if "offen" in m else np.zeros(64, np.uint64)
        off = off + np.uint64(self.sread(soff)) + np.uint64(int(m.get("offset", "0"), 0))
        return base, nrec, off

    def op_buffer_load_dwordx4(self, o, m):
        base, nrec, off = self._buf(o, m)
        F, i = self.vfile(o[0])
        for l in np.nonzero(self.execmask())[0]:
            for j in range(4):
                a = int(off[l]) + 4 * j
                F[i + j][l] = self.m.rd32(base + a) if a + 4 <= nrec else 0

    def _bufload(self, o, m, n):
        base, nrec, off = self._buf(o, m)
        F, i = self.vfile(o[0])
        for l in np.nonzero(self.execmask())[0]:
            for j in range(n):
                a = int(off[l]) + 4 * j
                F[i + j][l] = self.m.rd32(base + a) if a + 4 <= nrec else 0

    def op_buffer_load_dwordx2(self, o, m): self._bufload(o, m, 2)
    def op_buffer_load_dword(self, o, m): self._bufload(o, m, 1)
    def op_v_add_lshl_u32(self, o, m):
        a = (self.vread(o[1]).astype(np.uint64) + self.vread(o[2])) & MASK32
        self.vwrite(o[0], (a << (self.vread(o[3]) & 31).astype(np.uint64)) & MASK32)
    def op_buffer_load_dwordx3(self, o, m): self._bufload(o, m, 3)

    def op_buffer_store_dwordx4(self, o, m):
        base, nrec, off = self._buf(o, m)
        F, i = self.vfile(o[0])
        for l in np.nonzero(self.execmask())[0]:
            for j in range(4):
                a = int(off[l]) + 4 * j
                if a + 4 <= nrec:
                    self.m.wr32(base + a, int(F[i + j][l]))

    def op_global_atomic_add(self, o, m):
        addr = self.gaddr(o[1:2] + [o[3]], m) if False else None
        vaddr, vdata, saddr = o[1], o[2], o[3]
        a = self.vread(vaddr).astype(np.uint64) + np.uint64(self.s64(reg(saddr)[1]))
        a = a + np.uint64(int(m.get("offset", "0"), 0))
        e = self.execmask()
        k, i, nn = reg(o[0])
        data = self.vread(vdata)
        for l in np.nonzero(e)[0]:
            old = self.m.rd32(int(a[l]))
            self.m.wr32(int(a[l]), old + int(data[l]))
            if "sc0" in m:
                self.v[i][l] = old


def run_waves(m, prog, labels, kernarg, nwaves, budget, rng, quantum=(20, 200)):
    waves = [Wave(m, prog, labels, kernarg, w) for w in range(nwaves)]
    total = 0
    while any(not w.done for w in waves):
        live = [w for w in waves if not w.done]
        w = live[rng.integers(len(live))]
        q = rng.integers(*quantum)
        for _ in range(q):
            r = w.step()
            m.clock += 1
            total += 1
            if w.done or r == "yield":
                break
        if total > budget:
            return False, waves, total
    return True, waves, total
