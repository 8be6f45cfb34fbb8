"""Writes tests/golden/gem_consortium.npz: the three member models of the
reference's cooperative consortium (Actinoplanes sp. OR16, Rhizobacter
gummiphilus NS21, Propionibacterium freudenreichii shermanii; README.md of
the reference) parsed from the reference's SBML files with lpe.sbml, as
input data for tests that run where the SBML files are absent.

Run from the repository root:  python tests/golden/make_gem_fixture.py
"""

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "co-cultivation-metabolic-atlas_amd"))

from lpe import sbml  # noqa: E402

SBML_DIR = "/root/reference/models/sbml"
MEMBERS = ("final_consortium/Actinoplanes_sp_OR16_lcp.xml",
           "final_consortium/Rhizobacter_gummiphilus_NS21.xml",
           "helper_candidates/Propionibacterium_freudenreichii_shermanii_curated.xml")
OUT = os.path.join(ROOT, "tests", "golden", "gem_consortium.npz")

if __name__ == "__main__":
    gems = [sbml.read_sbml(os.path.join(SBML_DIR, m)) for m in MEMBERS]
    sbml.save_gems(OUT, gems)
    for g in gems:
        print(g.name, g.S.shape, g.S.nnz, len(g.exchanges))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
