"""Writes the src_rec table fixtures: the header and every row of a few events
of the reference's example observation tables (data files, copied row for
row, in file order):

  src_rec_checkerboard.csv  examples/00_checkerboard_iso/src_rec_file_ph.csv,
                            events MX00 and MX05 (16 periods, 23 receivers)
  src_rec_checkerboard_gr.csv  examples/00_checkerboard_iso/src_rec_file_gr.csv
                            (group-velocity times), the same two events
  src_rec_hawaii.csv        examples/04_hawaii_topo/src_rec_file_raw.csv,
                            its first 6 event names in byte order (15 periods,
                            uneven receiver counts; column order differs)

Run from the repository root with the reference tree at /root/reference.
"""
import os
import sys

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))


def extract(src, dst, pick):
    with open(src, "rb") as f:
        lines = f.read().split(b"\n")
    header = lines[0].decode().split(",")
    ie = header.index("evtname")
    body = [ln for ln in lines[1:] if ln]
    names = sorted({ln.split(b",")[ie] for ln in body})
    keep = set(pick(names))
    out = [lines[0]] + [ln for ln in body if ln.split(b",")[ie] in keep]
    with open(dst, "wb") as f:
        f.write(b"\n".join(out) + b"\n")
    print(dst, len(out) - 1, "rows", sorted(keep))


def main():
    if not os.path.isdir(REF):
        sys.exit("reference examples not found")
    extract(os.path.join(REF, "00_checkerboard_iso", "src_rec_file_ph.csv"),
            os.path.join(HERE, "src_rec_checkerboard.csv"), lambda n: [b"MX00", b"MX05"])
    extract(os.path.join(REF, "00_checkerboard_iso", "src_rec_file_gr.csv"),
            os.path.join(HERE, "src_rec_checkerboard_gr.csv"), lambda n: [b"MX00", b"MX05"])
    extract(os.path.join(REF, "04_hawaii_topo", "src_rec_file_raw.csv"),
            os.path.join(HERE, "src_rec_hawaii.csv"), lambda n: n[:6])


if __name__ == "__main__":
    main()
