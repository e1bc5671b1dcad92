"""Structure checks of the compiled gfx950 kernels (CPU only).

Two properties of the AMD assembly built from cpim_amd/csrc/engine.hip:

* A loop that contains s_barrier must leave through a uniform (scalar)
  branch.  If the compiler lowers it into an exec-mask loop
  (s_andn2_b64 exec, exec, ...; s_cbranch_execz), some lanes can keep
  iterating past a barrier while another lane of the same wave waits outside
  the loop: a workgroup that publishes its next work item from one thread then
  repeats the same item forever.
* A loop that sleeps (s_sleep, i.e. a spin-wait on another workgroup) must
  read the device wall clock (s_memrealtime / s_memtime) on every iteration:
  that is the bounded wait that turns a missing partner into an error flag
  instead of a hang.

Loops are recovered from the compiler's block comments ("Loop Header",
"Header=BBx_y", "Parent Loop BBx_y"): every block is assigned to every loop
that encloses it.
"""
import re


def _kernels(asm_text):
    kernels = {}
    cur = None
    for line in asm_text.split("\n"):
        m = re.match(r"^(_Z\w+):", line)
        if m:
            cur = m.group(1)
            kernels[cur] = []
        elif cur:
            kernels[cur].append(line)
            if "s_endpgm" in line:
                cur = None
    return kernels


def kernel_loops(asm_text):
    """{kernel: {loop header: [(block label, [lines])]}} over every loop."""
    result = {}
    for name, lines in _kernels(asm_text).items():
        blocks = []
        label, chain, body = ".entry", [], []
        for ln in lines:
            m = re.match(r"^(\.LBB\w+|; %bb\.\d+):\s*(;.*)?$", ln)
            if m:
                blocks.append((label, chain, body))
                label, body = m.group(1), []
                com = m.group(2) or ""
                chain = re.findall(r"Header=(BB\w+)", com) + re.findall(r"Parent Loop (BB\w+)", com)
                if "Loop Header" in com and "Header=" not in com:
                    chain.append(label[2:] if label.startswith(".L") else label)
            else:
                body.append(ln)
                chain = chain + re.findall(r"Parent Loop (BB\w+)", ln)
        blocks.append((label, chain, body))
        loops = {}
        for lab, ch, b in blocks:
            for h in set(ch):
                loops.setdefault(h, []).append((lab, b))
        result[name] = loops
    return result


def _has(blocks, pattern):
    return any(re.search(pattern, l) for _, b in blocks for l in b)


def barrier_loops(asm_text):
    """{kernel: sorted headers of the loops that contain s_barrier}."""
    return {k: sorted(h for h, bl in loops.items() if _has(bl, r"\bs_barrier\b"))
            for k, loops in kernel_loops(asm_text).items()}


def barrier_loops_with_exec_exit(asm_text):
    """{kernel: sorted headers of barrier loops left through an exec-mask branch}."""
    result = {}
    for name, loops in kernel_loops(asm_text).items():
        bad = set()
        for h, bl in loops.items():
            if not _has(bl, r"\bs_barrier\b"):
                continue
            labels = {lab[2:] if lab.startswith(".L") else lab for lab, _ in bl}
            for _, b in bl:
                for i, l in enumerate(b):
                    if "s_andn2_b64 exec, exec" in l:
                        m = re.search(r"s_cbranch_execz \.L(BB\w+)", " ".join(b[i + 1:i + 3]))
                        if m and m.group(1) not in labels:
                            bad.add(h)
        result[name] = sorted(bad)
    return result


def unbounded_spin_loops(asm_text):
    """{kernel: sorted headers of loops that sleep without reading the wall clock}.
    Only innermost sleeping loops matter: an outer loop around a bounded spin
    contains the clock read of that spin."""
    result = {}
    for name, loops in kernel_loops(asm_text).items():
        bad = []
        for h, bl in loops.items():
            if _has(bl, r"\bs_sleep\b") and not _has(bl, r"\bs_mem(real)?time\b"):
                bad.append(h)
        result[name] = sorted(bad)
    return result


def sleeping_loops(asm_text):
    """{kernel: sorted headers of the loops that contain s_sleep}."""
    return {k: sorted(h for h, bl in loops.items() if _has(bl, r"\bs_sleep\b"))
            for k, loops in kernel_loops(asm_text).items()}
