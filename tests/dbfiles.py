"""Writers for the reference's database and read file formats, used to drive the CLI
(DB.mtg/DB.mgs layout: src/util/ProgEnv.cpp:24-28, src/MetaGenome.cpp:42-73,
src/Genome.cpp:29-97)."""
import struct

import numpy as np

from msgseqtk_amd import synth


def _str(b):
    return struct.pack("<Q", len(b)) + b


def write_db(prefix, chroms, names):
    with open(prefix + ".mtg", "wb") as f:
        f.write(b"MSGseqTK" + struct.pack("<iii", 1, 2, 3))
        f.write(struct.pack("<Q", len(chroms)))          # one genome per chromosome
        for c, nm in zip(chroms, names):
            f.write(_str(nm.encode()) + _str(nm.encode()) + struct.pack("<Q", 1))
            f.write(_str(nm.encode()) + struct.pack("<q", len(c)))
    with open(prefix + ".mgs", "wb") as f:
        for c in chroms:
            f.write(np.asarray(c, np.uint8).tobytes() + b"\0")


def write_fastq(path, seqs, quals, names):
    with open(path, "w") as f:
        for s, q, n in zip(seqs, quals, names):
            f.write("@%s\n%s\n+\n%s\n" % (n, synth.decode(s), bytes(np.asarray(q, np.uint8) + 33).decode()))
