"""Copy the reference's demo input files (src/demo/) into tests/golden/demo/ as fixtures.

The files are data: a 20-particle point file written by SavePointsCSV2, the SimParams.txt
written by WriteSimParams and the genome.csv written by WriteGenome. genome.csv is cut after
its last data row (the tanh_param block): what follows in the reference's file is a
free-text description of the FGenome struct that ReadGenome never reads.

Run from the repository root in a container that has the reference checkout:
    python tests/golden/copy_demo_inputs.py /root/reference/Morphogenesis/src/demo
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ("SimParams.txt", "genome.csv", "particles_pos_vel_color100001.csv")


def genome_data_part(text: str) -> str:
    marker = "Bond remodelling m_FGenome.tanh_param[3][8]"
    k = text.index(marker)
    lines = text[k:].split("\n")
    keep = 3 + 3                     # title, two column-title lines, three rows
    return text[:k] + "\n".join(lines[:keep]) + "\n"


def main(src: str):
    dst = os.path.join(HERE, "demo")
    os.makedirs(dst, exist_ok=True)
    for name in FILES:
        with open(os.path.join(src, name), "r") as fh:
            text = fh.read()
        if name == "genome.csv":
            text = genome_data_part(text)
        with open(os.path.join(dst, name), "w") as fh:
            fh.write(text)
        print(name, len(text), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Morphogenesis/src/demo")
