"""Write xfluids_amd/data/runtime.dat: the species thermo fits and mechanisms
the benchmark workloads use, in the reference's runtime.dat file formats.

Reads the reference tree (default /root/reference/runtime.dat); run once in
the build container.  Copies the numeric blocks of the species it needs from
thermal_dynamics.dat (NASA-9) and thermal_dynamics_janaf.dat (NASA-7), the
H2/O2 19-reaction mechanisms and the transport tables (transport_data.dat,
collision_integral.dat), and writes the mixture lists of this project's own
workloads (air/H2 shock-bubble, H2/N2 vs air mixing layer) and copies the
reference's own shock-bubble mixtures (Inert-SBI, Inert-SBI-without-fuel,
RSBI-19REA) and the 21-reaction H2O-N2 mechanism.

    python tests/golden/make_runtime_subset.py [reference_runtime_dat]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(os.path.dirname(HERE)), "xfluids_amd", "data", "runtime.dat")
SPECIES = ["NCOP", "H", "O", "H2O", "OH", "O2", "H2", "H2O2", "HO2", "N2", "AR", "Xe"]


def blocks(path):
    """Map '*NAME' -> list of following lines up to the next '*' keyword."""
    out, cur = {}, None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("*"):
                cur = s.split()[0]
                if cur == "*END":
                    break
                out[cur] = []
            elif cur is not None and s and not s.startswith("!") and not s.startswith("---"):
                out[cur].append(line.rstrip("\n"))
    return out


def write_thermo(src_root):
    os.makedirs(OUT, exist_ok=True)
    for fname, title in (("thermal_dynamics.dat", "NASA-9 (3 ranges): a1-a7 b1 b2 per range, then W [g/mol]"),
                         ("thermal_dynamics_janaf.dat", "NASA-7 (JANAF): a1-a7 high, a1-a7 low, then W [g/mol]")):
        b = blocks(os.path.join(src_root, fname))
        with open(os.path.join(OUT, fname), "w") as f:
            f.write("! subset of the XFluids runtime.dat file of the same name: %s\n" % title)
            for sp in SPECIES:
                key = "*" + sp
                if key not in b:
                    continue
                f.write(key + "\n")
                for line in b[key]:
                    f.write(line + "\n")
            f.write("*END\n")


def copy_mech(src_root, name):
    d = os.path.join(OUT, "Reaction", name)
    os.makedirs(d, exist_ok=True)
    for fn in ("species_list.dat", "reaction_list.dat"):
        with open(os.path.join(src_root, "Reaction", name, fn)) as fi, open(os.path.join(d, fn), "w") as fo:
            fo.write(fi.read())


def copy_mixture(src_root, name):
    d = os.path.join(OUT, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(src_root, name, "species_list.dat")) as fi, \
            open(os.path.join(d, "species_list.dat"), "w") as fo:
        fo.write(fi.read())


def copy_root_file(src_root, fn):
    with open(os.path.join(src_root, fn)) as fi, open(os.path.join(OUT, fn), "w") as fo:
        fo.write(fi.read())


def write_mixture(rel, names, out_ratio, in_ratio):
    d = os.path.join(OUT, rel)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "species_list.dat"), "w") as f:
        f.write("\t".join(names) + "\n")
        f.write("\t".join(str(v) for v in out_ratio) + "\n")
        f.write("\t".join(str(v) for v in in_ratio) + "\n")


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/runtime.dat"
    write_thermo(src)
    write_mixture("NO-COP", ["NCOP"], [1], [1])
    # air outside (O2:N2 = 21:79 molar), pure H2 inside the bubble; N2 last
    write_mixture("H2-O2-N2", ["H2", "O2", "N2"], [0, 21, 79], [1, 0, 0])
    # two-species variant of the same bubble (N2 outside, H2 inside): the
    # NS = 2 kernel instances
    write_mixture("H2-N2", ["H2", "N2"], [0, 1], [1, 0])
    copy_mech(src, "H2O_19_reaction")
    copy_mech(src, "H2O-N2_19_reaction")
    # Lennard-Jones parameters and collision integrals for the transport fits
    copy_root_file(src, "transport_data.dat")
    copy_root_file(src, "collision_integral.dat")
    # mixing layer: air outside (upper stream), H2:N2 = 1:1 inside (lower stream)
    write_mixture("Reaction/H2O-N2_19_reaction", ["H2", "O2", "O", "H", "OH", "HO2", "H2O2", "H2O", "N2"],
                  [0, 21, 0, 0, 0, 0, 0, 0, 79], [1, 0, 0, 0, 0, 0, 0, 0, 1])
    # the reference's own shock-bubble mixtures, whose last species is a
    # GhostSpecies copy of N2: Inert-SBI (H2/O2/Xe bubble in N2, NS = 5),
    # Inert-SBI-without-fuel (O2/Xe bubble, NS = 4), and the reacting
    # RSBI-19REA (NS = 11, 19 reactions, Xe in place of AR)
    copy_mixture(src, "Inert-SBI")
    copy_mixture(src, "Inert-SBI-without-fuel")
    copy_mech(src, "RSBI-19REA")
    # H2O-N2_21_reaction (NS = 10 with the ghost N2, 21 reactions): the same
    # non-premixed streams as the 19-reaction mixing layer
    copy_mech(src, "H2O-N2_21_reaction")
    write_mixture("Reaction/H2O-N2_21_reaction", ["H2", "O2", "O", "H", "OH", "HO2", "H2O2", "H2O", "N2", "N2"],
                  [0, 21, 0, 0, 0, 0, 0, 0, 79, 0], [1, 0, 0, 0, 0, 0, 0, 0, 1, 0])
    print("wrote", OUT)


if __name__ == "__main__":
    main()
