"""Reader for the `.mol` basis/geometry files used by simint's tests.

Format (skel/test/Common.cpp:83-174): natom; per atom a header
``SYM nshell nallprim nallprimg`` and a line ``x y z``; per shell
``TYPE nprim ngen`` then nprim lines ``alpha c_1 .. c_ngen``.  A TYPE of
several letters (``SP``) is a general contraction sharing exponents; the
letters cycle over the ngen coefficient columns.
"""

_AM = {c: i for i, c in enumerate("SPDFGHIJKLMNOQRTUVWXYZABCE")}


def read_mol(path):
    """Return (shells, atoms): shells = [(am, (x,y,z), alphas, coefs)] in file
    order grouped as the reference's ShellMap (by AM, file order within an AM);
    atoms = [(symbol, x, y, z)]."""
    toks = open(path).read().split()
    pos = 0

    def nxt():
        nonlocal pos
        pos += 1
        return toks[pos - 1]

    natom = int(nxt())
    by_am = {}
    atoms = []
    for _ in range(natom):
        sym = nxt()
        nshell = int(nxt()); nxt(); nxt()
        x, y, z = float(nxt()), float(nxt()), float(nxt())
        atoms.append((sym, x, y, z))
        for _ in range(nshell):
            typ = nxt()
            nprim = int(nxt())
            ngen = int(nxt())
            alphas = []
            coefs = [[] for _ in range(ngen)]
            for _ in range(nprim):
                alphas.append(float(nxt()))
                for g in range(ngen):
                    coefs[g].append(float(nxt()))
            for g in range(ngen):
                am = _AM[typ[g % len(typ)]]
                by_am.setdefault(am, []).append((am, (x, y, z), list(alphas), coefs[g]))
    shells = []
    for am in sorted(by_am):
        shells.extend(by_am[am])
    return shells, atoms
