"""Writes chain_asm.h: sequential FP64 sums of n doubles in LDS (n = 512, 64, 16) as one inline
assembly block each.

A sequential sum is bound by its dependent FP64 adds (5.7 clocks each on MI355X, measured in a
register-only chain) and by the instructions that bring the values in: a single wave pays
about 11 clocks to issue one LDS read and 5 clocks per s_waitcnt.  Measured per add, one wave,
eight chains in eight lanes: the compiler's own code (two or three 128-bit reads ahead) ~16;
64-bit reads with a wait per value 16.3; 128-bit reads with a wait per read 10.4; 128-bit
reads waited for and issued in groups of GROUP, INFLIGHT outstanding, 8.1.

Each block keeps INFLIGHT 128-bit reads outstanding in a ring of INFLIGHT + GROUP register
quads at fixed registers v[BASE : BASE + 4 * (INFLIGHT + GROUP) - 1] (declared clobbered, so
the compiler keeps nothing there across the block).  A step waits for the oldest GROUP reads,
issues the next GROUP reads, then adds the 2 * GROUP values in order; every read is waited for
inside the block (the last step waits for all), so no load is in flight when the block ends.
BASE is a template argument: 16 for kernels that chain at the end of a short body, 200 in
k_bottom, whose Krylov state occupies the low registers.

A second form, seg_chain512_asm<BASE>(addr, init), runs two 512-term sums in one wave with every
value in registers: lane 32 h + j (h = 0, 1) reads values 16 j .. 16 j + 15 of sum h (eight
128-bit reads, all waited for before the first add), and the running sum moves one lane to the
right after each lane's 16 adds (DPP wave_shr:1 on both halves of the double; every lane adds
its own values at every step, so no EXEC changes are needed, and lane 32 h + s holds sum h up
to its segment after step s).  Measured on MI355X, including the reads: 6.3 clocks per add
against 7.85 for the 128-bit ring (5.68 is a register-only chain).  s_nop 1 covers the two
wait states a DPP read needs after a VALU write of the same register.

    python sayakaamr_amd/csrc/dev/gen_chain_asm.py   # rewrites chain_asm.h beside this script
"""
import os

INFLIGHT = 8
GROUP = 4
RING = INFLIGHT + GROUP
SIZES = (512, 64, 16)
BASES = (16, 200)


def block(n, base):
    nr = n // 2                       # 128-bit reads
    quad = lambda j: f"v[{base + 4 * (j % RING)}:{base + 4 * (j % RING) + 3}]"
    lo = lambda j: f"v[{base + 4 * (j % RING)}:{base + 4 * (j % RING) + 1}]"
    hi = lambda j: f"v[{base + 4 * (j % RING) + 2}:{base + 4 * (j % RING) + 3}]"
    lines = ["s_waitcnt lgkmcnt(0)"]
    issued = 0
    while issued < min(INFLIGHT, nr):
        lines.append(f"ds_read_b128 {quad(issued)}, %[a] offset:{16 * issued}")
        issued += 1
    i = 0
    while i < nr:
        g = min(GROUP, nr - i)
        lines.append(f"s_waitcnt lgkmcnt({issued - (i + g)})")   # reads i .. i+g-1 have landed
        for _ in range(g):
            if issued < nr:
                lines.append(f"ds_read_b128 {quad(issued)}, %[a] offset:{16 * issued}")
                issued += 1
        for j in range(i, i + g):
            lines.append(f"v_add_f64 %[s], %[s], {lo(j)}")
            lines.append(f"v_add_f64 %[s], %[s], {hi(j)}")
        i += g
    return lines


SEG = 16           # values per lane in seg_chain512_asm (512 = 32 lanes x 16)
SEG_BASES = (16, 200)


def seg_block(base):
    acc = base + 2 * SEG
    lines = ["s_waitcnt lgkmcnt(0)"]
    for k in range(SEG // 2):
        lines.append(f"ds_read_b128 v[{base + 4 * k}:{base + 4 * k + 3}], %[a] offset:{16 * k}")
    lines += [f"v_mov_b32 v{acc}, %[ilo]", f"v_mov_b32 v{acc + 1}, %[ihi]", "s_waitcnt lgkmcnt(0)"]
    for step in range(512 // SEG):
        if step:
            lines.append("s_nop 1")
            lines.append(f"v_mov_b32_dpp v{acc}, v{acc} wave_shr:1 row_mask:0xf bank_mask:0xf")
            lines.append(f"v_mov_b32_dpp v{acc + 1}, v{acc + 1} wave_shr:1 row_mask:0xf bank_mask:0xf")
        for k in range(SEG):
            lines.append(f"v_add_f64 v[{acc}:{acc + 1}], v[{acc}:{acc + 1}], v[{base + 2 * k}:{base + 2 * k + 1}]")
    lines += [f"v_mov_b32 %[lo], v{acc}", f"v_mov_b32 %[hi], v{acc + 1}"]
    return lines


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    out = [
        "// chain_asm.h -- generated by gen_chain_asm.py (same folder); do not edit by hand.",
        "//",
        "// lds_chain_asm<N, BASE>(addr, s): s + v[0] + v[1] + ... + v[N-1], added one at a time in",
        "// index order (the same sequence of roundings as a C loop), the N doubles at LDS byte address",
        f"// addr (16-byte aligned).  {INFLIGHT} 128-bit LDS reads stay in flight in a ring of register",
        f"// quads at v[BASE .. BASE + {4 * RING - 1}], waited for and reissued {GROUP} at a time; every read",
        "// is waited for inside the block.",
        "#pragma once",
        "",
        "namespace sydev {",
        "",
        "template <int N, int BASE>",
        "__device__ inline double lds_chain_asm(unsigned addr, double s);",
        "",
    ]
    for base in BASES:
        clob = ", ".join(f'"v{r}"' for r in range(base, base + 4 * RING))
        for n in SIZES:
            out.append("template <>")
            out.append(f"__device__ inline double lds_chain_asm<{n}, {base}>(unsigned addr, double s) {{")
            out.append("    asm volatile(")
            for ln in block(n, base):
                out.append(f'        "{ln}\\n"')
            out.append('        : [s] "+v"(s)')
            out.append('        : [a] "v"(addr)')
            out.append(f'        : "memory", {clob});')
            out.append("    return s;")
            out.append("}")
            out.append("")
    out += [
        "// seg_chain512_asm<BASE>(addr, init): two sequential 512-term sums per wave (whole wave",
        "// active), lane 32 h + j holding values 16 j .. 16 j + 15 of sum h at LDS byte address addr,",
        "// sum h starting from lane 32 h's init; the sum moves one lane right (DPP wave_shr:1) after",
        f"// each lane's {SEG} adds.  Result: sum 0 in lane 31, sum 1 in lane 63.  Registers",
        f"// v[BASE .. BASE + {2 * SEG + 1}] are used and clobbered.",
        "template <int BASE>",
        "__device__ inline double seg_chain512_asm(unsigned addr, double init);",
        "",
    ]
    for base in SEG_BASES:
        clob = ", ".join(f'"v{r}"' for r in range(base, base + 2 * SEG + 2))
        out.append("template <>")
        out.append(f"__device__ inline double seg_chain512_asm<{base}>(unsigned addr, double init) {{")
        out.append("    unsigned lo, hi;")
        out.append("    const unsigned ilo = (unsigned)__double2loint(init), ihi = (unsigned)__double2hiint(init);")
        out.append("    asm volatile(")
        for ln in seg_block(base):
            out.append(f'        "{ln}\\n"')
        out.append('        : [lo] "=v"(lo), [hi] "=v"(hi)')
        out.append('        : [a] "v"(addr), [ilo] "v"(ilo), [ihi] "v"(ihi)')
        out.append(f'        : "memory", {clob});')
        out.append("    return __hiloint2double((int)hi, (int)lo);")
        out.append("}")
        out.append("")
    out.append("}  // namespace sydev")
    with open(os.path.join(here, "chain_asm.h"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
