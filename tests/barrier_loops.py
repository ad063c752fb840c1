"""Static check of gfx950 assembly built from this package's kernels: no loop
whose exit is a per-lane exec mask may hold a workgroup barrier.

Such a loop lets some lanes of a wavefront wait at the loop exit while the
other lanes keep executing the barrier.  Workgroup-uniform source code can
still compile to one: the compiler may merge a "thread 0 does X" block at the
end of a persistent loop with one at its head and split the loop in two, so
that thread 0 waits outside while its wavefront mates spin inside on stale
shared data.  Usage: python barrier_loops.py file.s [...] (exit 1 if flagged).
"""
import re
import sys


def scan(text):
    """[(function, first body line, latch line)] of flagged loops."""
    lines = text.split("\n")
    labels = {}
    funcs = []
    for i, line in enumerate(lines):
        m = re.match(r"^(\.LBB\d+_\d+):", line)
        if m:
            labels[m.group(1)] = i
        elif re.match(r"^_Z\w+:", line):
            funcs.append((i, line.split(":")[0]))

    def func_of(i):
        name = None
        for start, n in funcs:
            if start <= i:
                name = n
        return name

    found = []
    for i, line in enumerate(lines):
        if "s_andn2_b64 exec, exec" not in line:
            continue
        j = i + 1
        while lines[j].strip().startswith((";", "s_nop")):
            j += 1
        nxt = lines[j].strip()
        m = re.match(r"s_cbranch_execnz (\.LBB\d+_\d+)", nxt)
        if m and labels[m.group(1)] < i:  # back edge upwards: body is head .. latch
            head = labels[m.group(1)]
            if any("s_barrier" in x for x in lines[head:i]):
                found.append((func_of(i), head + 1, i + 1))
            continue
        m = re.match(r"s_cbranch_execz (\.LBB\d+_\d+)", nxt)
        if m and j + 1 < len(lines) and re.match(r"^\.LBB\d+_\d+:", lines[j + 1]) \
                and "Loop Header" in "".join(lines[j + 1:j + 4]):
            # latch placed before the header: the loop runs down to its exit label
            end = labels[m.group(1)]
            if any("s_barrier" in x for x in lines[j + 1:end]):
                found.append((func_of(i), j + 2, i + 1))
    return found


def main(paths):
    bad = 0
    for p in paths:
        for fn, body, latch in scan(open(p).read()):
            print(f"{p}: {fn}: lane-divergent loop (body line {body}, latch line {latch}) holds s_barrier")
            bad += 1
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
