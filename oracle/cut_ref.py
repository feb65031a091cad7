"""Test infrastructure only (oracle/Makefile.ref): writes the text of named
external subroutines of a reference source file to another file, so that a
routine whose file cannot be compiled whole (ns2DComp.ALE.f90 holds the main
program, which uses smoothing_mod) can be compiled on its own.  The output
goes to a temporary directory and is deleted after compiling; only the
object lands in oracle/_ref.

    python oracle/cut_ref.py SRC OUT NAME [NAME ...]
"""
import re
import sys


def cut(lines, name):
    start = re.compile(r"\s*subroutine\s+" + name + r"\b", re.I)
    # a subroutine ends at `end subroutine NAME` or at a bare `end`
    end = re.compile(r"\s*end(\s+subroutine(\s+" + name + r")?)?\s*(!.*)?$", re.I)
    lo = next(i for i, l in enumerate(lines) if start.match(l))
    hi = next(i for i in range(lo + 1, len(lines)) if end.match(lines[i]))
    return lines[lo:hi + 1]


def main(argv):
    src, out, names = argv[1], argv[2], argv[3:]
    lines = open(src).read().splitlines(keepends=True)
    with open(out, "w") as fh:
        if lines and lines[0].startswith("#define"):
            fh.write(lines[0])         # the timer macro, should a routine use it
        for n in names:
            fh.writelines(cut(lines, n))


if __name__ == "__main__":
    main(sys.argv)
