"""Site database for the molecules the 3D-RISM path is exercised with.

Values are the reference's data table (guile/solutes.scm) for the entries
that its README, test/run-ions.scm and test/forces.scm use; a data table has
to be identical to give identical results.  Each site is
``(name, (x, y, z), sigma [A], epsilon [kcal/mol], charge [e])``, the order
of the reference's ``Site`` struct (bgy3d-solutes.h:9-16).

Unit forms are expanded as guile/molecule.scm:84-101 does: ``(r0 x)`` is a
Rmin/2 radius and becomes sigma = x * 2^(5/6).
"""

_R0 = 2.0 ** (5.0 / 6.0)


def _ion(name, r0, eps, q):
    return ((name, (0.0, 0.0, 0.0), r0 * _R0, eps, q),)


_X = 0.288675134594813
_Y = 0.816496580927726

DATABASE = {
    # guile/solutes.scm:307-320
    "water, SPC/E": (
        ("O", (-_X, 0.0, 0.0), 3.1656, 0.1553, -0.8476),
        ("OH", (_X, _Y, 0.0), 1.0, 0.0545, 0.4238),
        ("OH", (_X, -_Y, 0.0), 1.0, 0.0545, 0.4238)),
    "water, cSPC/E": (
        ("O", (-_X, 0.0, 0.0), 3.1656, 0.1553, -0.8476),
        ("OH", (_X, _Y, 0.0), 1.1658, 0.01553, 0.4238),
        ("OH", (_X, -_Y, 0.0), 1.1658, 0.01553, 0.4238)),
    "water, PR-SPC/E": (
        ("O", (-_X, 0.0, 0.0), 3.1656, 0.1553, -0.8476),
        ("OH", (_X, _Y, 0.0), 0.4, 0.046, 0.4238),
        ("OH", (_X, -_Y, 0.0), 0.4, 0.046, 0.4238)),
    # guile/solutes.scm:252-255 (test-suite solvent "water")
    "water": (
        ("O", (-0.2929, 0.0, 0.0), 3.1506, 0.1521, -0.834),
        ("OH", (0.2929, 0.757, 0.0), 0.4, 0.046, 0.417),
        ("OH", (0.2929, -0.757, 0.0), 0.4, 0.046, 0.417)),
    # guile/solutes.scm:346-352
    "OW": (("OW", (0.0, 0.0, 0.0), 3.16, 0.1549, 0.0),),
    "OW2": (("OW", (0.0, 0.0, 2.0), 3.16, 0.1549, 0.0),
            ("OW", (0.0, 0.0, -2.0), 3.16, 0.1549, 0.0)),
    "OW2-UA": (("UA", (0.0, 0.0, 0.0), 3.16, 0.6196, 0.0),),
    # guile/solutes.scm:597-605, JC08 ions for SPC/E
    "Li+, JC08, SPC/E": _ion("Li", 0.791, 0.3367344, 1.0),
    "Na+, JC08, SPC/E": _ion("Na", 1.212, 0.3526418, 1.0),
    "K+, JC08, SPC/E": _ion("K", 1.593, 0.4297054, 1.0),
    "Rb+, JC08, SPC/E": _ion("Rb", 1.737, 0.4451036, 1.0),
    "Cs+, JC08, SPC/E": _ion("Cs", 2.021, 0.0898565, 1.0),
    "F-, JC08, SPC/E": _ion("F", 2.257, 0.0074005, -1.0),
    "Cl-, JC08, SPC/E": _ion("Cl", 2.711, 0.0127850, -1.0),
    "Br-, JC08, SPC/E": _ion("Br", 2.751, 0.0269586, -1.0),
    "I-, JC08, SPC/E": _ion("I", 2.919, 0.0427845, -1.0),
}


def find_molecule(name):
    """Sites of a named molecule (guile/molecule.scm:148 find-molecule)."""
    try:
        return DATABASE[name]
    except KeyError:
        raise KeyError(f"no such molecule: {name!r}") from None
