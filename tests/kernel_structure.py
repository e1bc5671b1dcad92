"""Structure check of the compiled gfx950 kernels (CPU only).

A loop that contains s_barrier must leave through a uniform (scalar) branch.
If the compiler lowers it into an exec-mask loop (s_andn2_b64 exec, exec, ...;
s_cbranch_execz), some lanes can keep iterating past a barrier while another
lane of the same wave waits outside the loop: a workgroup that publishes its
next work item from one thread then repeats the same item forever.  This
module parses the AMD assembly built from cpim_amd/csrc/engine.hip and lists
such loops per kernel.
"""
import re


def barrier_loops_with_exec_exit(asm_text):
    text = asm_text.split("\n")
    kernels = {}
    cur = None
    for line in text:
        m = re.match(r"^(_Z\w+):", line)
        if m:
            cur = m.group(1)
            kernels[cur] = []
        elif cur:
            kernels[cur].append(line)
            if "s_endpgm" in line:
                cur = None
    result = {}
    for name, lines in kernels.items():
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
        bad = set()
        for h, bl in loops.items():
            if not any("s_barrier" in l for _, b in bl for l in b):
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
