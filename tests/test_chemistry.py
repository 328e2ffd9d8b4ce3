"""Chemistry decks in the reference's format (read_reactions,
mod_reactions.f90:13-411) and the tables built from them (ordering,
:414-558).  CPU only."""
import os

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import chemistry as CH
from essaitemporaire_lepic_module_amd import config as cfg

REF_DECK = "/root/reference/input_dir/argon.dat"
NGAS = 1.0 / (cfg.KB * 300.0)


def _synthetic_reactions():
    E = cfg.argon_energy_grid(200)
    rs = [((0, 2), (0, 2), 0.0, 1, np.full_like(E, 1e-19)),
          ((0, 2), (0, 2), 11.5, 3, cfg.lotz_like(E, cfg.S0_EXCITATION, 11.5)),
          ((0, 2), (0, 1, 0), 15.8, 2, cfg.lotz_like(E, cfg.S0_IONIZATION, 15.8)),
          ((1, 2), (1, 2), 0.0, 4, np.full_like(E, 5e-19))]
    return [CH.Reaction("", re, by, Eth, 0.0, kind, sum(1 for p in by if p == 0), E.copy(), sg)
            for re, by, Eth, kind, sg in rs]


def test_deck_round_trip_gives_the_synthetic_tables():
    """cfg2's analytic argon set written as a deck and read back gives the
    tables built from it directly, to the bit."""
    sp = cfg.argon_species()
    reactions = _synthetic_reactions()
    text = CH.write_reactions(sp, [1.0, 1.0, 1.0], reactions)
    chem = CH.read_reactions(text, NGAS)
    assert [s.name for s in chem.species] == ["e", "Ar+", "Ar"]
    assert [s.tracked for s in chem.species] == [True, True, False]
    for a, b in zip(chem.species[1:], sp[1:]):
        assert a.mass == pytest.approx(b.mass, rel=1e-15) and a.charge == b.charge
    got = CH.build_tables(chem, NGAS)
    want = cfg.argon_chemistry(sp, NGAS)
    for k in ("sig_Er", "sig", "sig_list", "p_ncol", "sig_Eth", "ncol", "npt"):
        np.testing.assert_array_equal(np.asarray(got[k]), np.asarray(want[k]), err_msg=k)
    np.testing.assert_allclose(got["sigv_mx"], want["sigv_mx"], rtol=1e-15)
    ci = np.asarray(got["col_info"]).reshape(-1, 10)
    np.testing.assert_array_equal(ci[:, :2], [[2, 2], [2, 2], [2, 3], [2, 2]])
    # reaction kind and product-electron count after the species (mod_reactions.f90:255-273)
    np.testing.assert_array_equal(ci[2, 7:9], [2, 2])
    np.testing.assert_array_equal(ci[3, 6:8], [4, 0])


FEATURES = """
comment lines before the sections are skipped
neut
name   mass  charge  temp.(eV)  n/n0
-----------------------------
[He]     4.0     0.     0.05    0.5
-----------------------------
Ions
name   mass  charge  temp.(eV)  n/n0
-----------------------------
[He+]    4.0     1.     0.1     1.
[H-]     1.0    -1.     0.1     0.2
-----------------------------
SURF
name ...
--------
[He+]  [He]  0.5  0.
--------
reac
[e]+[He]->[e]+[He+]+[e]
 24.6  0.
 1.  1.e-20
ionization
----------
  0.   0.
  24.6 0.
  100. 3.
----------
"""


def test_deck_features():
    """Lower-case keywords, a neutral defined before the ions (renumbered so
    that charged species come first), a surface section, scale factors and
    the '[H-]' tag."""
    chem = CH.read_reactions(FEATURES, 2.0e20)
    assert [s.name for s in chem.species] == ["e", "He+", "H-", "He"]
    assert chem.density == [1.0, 1.0, 0.2, 0.5 * 2.0e20]
    assert chem.tag_neg == 2
    assert chem.species[3].mass == pytest.approx(-4.0 * cfg.AMU)
    assert chem.species[2].charge == -cfg.QE
    (r,) = chem.reactions
    assert r.reactants == (0, 3) and r.products == (0, 1, 0)
    assert r.kind == CH.KINDS["IONIZATION"] and r.n_electrons == 2
    assert r.Eth == 24.6
    np.testing.assert_array_equal(r.E, [0.0, 24.6, 100.0])
    np.testing.assert_array_equal(r.sigma, [0.0, 0.0, 3.0 * 1.e-20])   # value times scale


def test_deck_without_gas_has_no_collisions():
    assert CH.read_reactions(FEATURES, 0.0).reactions == []


def test_deck_errors():
    with pytest.raises(CH.DeckError):
        CH.read_reactions("REACTION\n[e]+[X]->[e]+[X]\n", 1e20)
    bad = FEATURES.replace("[e]+[He]->[e]+[He+]+[e]", "[e]+[Xe]->[e]+[Xe+]")
    with pytest.raises(CH.DeckError):
        CH.read_reactions(bad, 1e20)
    with pytest.raises(CH.DeckError):
        CH.read_reactions(FEATURES.replace("ionization", "fusion"), 1e20)


@pytest.mark.skipif(not os.path.exists(REF_DECK), reason="the reference's deck is not here")
def test_reference_argon_deck():
    """The reference's own argon deck: reaction count, table lengths and the
    cross-section scale factors, counted from the file by a plain scan."""
    lines = open(REF_DECK).read().splitlines()
    heads = [i for i, l in enumerate(lines) if l.strip() == "REACTION"]
    chem = CH.read_reactions(REF_DECK, NGAS)
    assert len(chem.reactions) == len(heads) == 6
    assert [s.name for s in chem.species] == ["e", "Ar+", "Ar"]
    for h, r in zip(heads, chem.reactions):
        assert r.label == lines[h + 1].strip()
        s_E, s_sig = (float(v) for v in lines[h + 3].split())
        first = h + 6                     # label, Eth/dE, scales, kind, dashes
        assert lines[first - 1].startswith("---")
        end = next(i for i in range(first, len(lines)) if lines[i].startswith("---"))
        raw = np.array([[float(v) for v in lines[i].split()] for i in range(first, end)])
        assert len(r.E) == end - first
        np.testing.assert_array_equal(r.E, raw[:, 0] * s_E)
        np.testing.assert_array_equal(r.sigma, raw[:, 1] * s_sig)
    t = CH.build_tables(chem, NGAS)
    Er = np.asarray(t["sig_Er"])
    assert (np.diff(Er) > 0).all()
    np.testing.assert_array_equal(Er, np.unique(np.concatenate([r.E for r in chem.reactions])))
    np.testing.assert_array_equal(t["p_ncol"], [4, 2, 0])
