"""Chemistry decks in the reference's format.

read_reactions restates the parser of modules/particles/mod_reactions.f90:13-411
(the deck's IONS / NEUTRALS / REACTION / SURF sections, read with Fortran
list-directed and A-edit semantics); build_tables hands the result to the
restatement of `ordering` (mod_reactions.f90:414-558) in
config.build_reaction_tables.  write_reactions writes a chemistry back in the
same format, so that synthetic cross sections go through the same path as a
user's deck.

Species are numbered as the reference numbers them: the electron first
(mod_reactions.f90:103-110), then every ion and neutral in the order the deck
defines them.  The device kernels need the charged (tracked) species before
the neutral ones; a deck that defines a neutral before an ion is renumbered
so that they are, and the reactions are renumbered with it.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from . import config as cfg

# reaction kinds (mod_reactions.f90:255-259)
KINDS = {"COLLISION": 1, "IONIZATION": 2, "EXCITATION": 3, "CHARGEEXCHANGE": 4,
         "DISSOCIATION": 5}
KIND_NAMES = {v: k for k, v in KINDS.items()}


class DeckError(ValueError):
    """A deck the reference's parser stops on (stop_calculation)."""


@dataclasses.dataclass
class Reaction:
    label: str
    reactants: tuple      # species indices (0-based)
    products: tuple
    Eth: float            # eV, scaled by the energy factor
    dE: float
    kind: int             # KINDS
    n_electrons: int      # electrons among the products
    E: np.ndarray         # eV, scaled
    sigma: np.ndarray     # m^2, scaled


@dataclasses.dataclass
class Chemistry:
    species: list         # config.Species, the electron first
    density: list         # n/n0 for charged species, n/n0 * ngas for neutrals
    reactions: list
    tag_neg: int = -1     # '[H-]', '[H]', '[eb]' (0-based, -1: absent)
    tag_neu: int = -1
    tag_beam: int = -1


def _first_token(line: str) -> str:
    """The first value of a list-directed read of a character variable:
    up to the first blank, comma or slash."""
    s = line.lstrip()
    for i, ch in enumerate(s):
        if ch in " \t,/":
            return s[:i]
    return s


class _Records:
    """Record-by-record reader with the two kinds of reads the parser uses."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.i = 0

    def eof(self) -> bool:
        return self.i >= len(self.lines)

    def list_token(self, width: int) -> str:
        """read(10,*) name with character(len=width): blank records are
        skipped, the first value is truncated to the variable's length."""
        while not self.eof():
            line = self.lines[self.i]
            self.i += 1
            tok = _first_token(line)
            if tok:
                return tok[:width]
        raise EOFError

    def list_values(self, count: int, first: str | None = None) -> list:
        """read(10,*) of count numbers, across records if needed; first is
        the rest of a record a non-advancing read left partly consumed."""
        vals = []
        pending = first
        while len(vals) < count:
            if pending is None:
                if self.eof():
                    raise EOFError
                pending = self.lines[self.i]
                self.i += 1
            for tok in pending.replace(",", " ").split():
                if tok.startswith("/"):
                    break
                try:
                    vals.append(float(tok.replace("d", "e").replace("D", "e")))
                except ValueError:
                    raise DeckError(f"not a number: {tok!r}") from None
                if len(vals) == count:
                    break
            pending = None
        return vals

    def record(self) -> str:
        """A formatted read of the next record (blank records are records)."""
        if self.eof():
            raise EOFError
        line = self.lines[self.i]
        self.i += 1
        return line


def read_reactions(deck, ngas: float, Te: float = 3.0) -> Chemistry:
    """Parse a chemistry deck (a path or its text) as read_reactions
    (mod_reactions.f90:13-411) does.  ngas scales the neutral densities; Te
    is the electron temperature of the run (the deck does not hold it)."""
    if isinstance(deck, (str, os.PathLike)) and os.path.exists(deck):
        with open(deck) as f:
            text = f.read()
    else:
        text = str(deck)
    rec = _Records(text)
    # the electron (mod_reactions.f90:103-110)
    names = ["[e]"]
    species = [cfg.Species("e", cfg.ME, -cfg.QE, Te)]
    density = [1.0]
    chem = Chemistry(species, density, [])
    seen_particles = False
    try:
        while True:
            name = rec.list_token(4)
            if name in ("IONS", "Ions", "ions", "NEUT", "Neut", "neut"):
                seen_particles = True
                _read_species(rec, name, names, chem, ngas)
            elif name in ("REAC", "Reac", "reac"):
                if not seen_particles:
                    break
                chem.reactions.append(_read_reaction(rec, names))
            elif name in ("SURF", "Surf", "surf"):
                if not seen_particles:
                    break
                _skip_surface(rec)
    except EOFError:
        pass
    if not seen_particles:
        raise DeckError('the particle sections must come first')
    if ngas == 0.0:
        chem.reactions = []          # collision-less mode (mod_reactions.f90:378)
    return _tracked_first(chem)


def _read_species(rec: _Records, name: str, names: list, chem: Chemistry, ngas: float):
    while name != "----":
        name = rec.list_token(4)
    while True:
        line = rec.record()
        tname = line[:6]
        if tname[:2] == "--":
            return
        pname = tname.rstrip()
        mass, charge, Ti, ni0 = rec.list_values(4, first=line[6:])
        if ni0 > 1.0:
            raise DeckError(f"(n/n0) > 1 for {pname}")
        charge = charge * cfg.QE
        mass = mass * cfg.AMU
        tracked = charge != 0.0
        if not tracked:
            ni0 = ni0 * ngas
            mass = -abs(mass)
        idx = len(names)
        names.append(pname)
        chem.species.append(cfg.Species(pname.strip("[]"), mass, charge, Ti, tracked=tracked))
        chem.density.append(ni0)
        if pname == "[H-]":
            chem.tag_neg = idx
        if pname == "[H]":
            chem.tag_neu = idx
        if pname == "[eb]":
            chem.tag_beam = idx


def _read_reaction(rec: _Records, names: list) -> Reaction:
    rlabel = rec.list_token(50)
    ids, n_re = [], None
    l_sy = r_sy = -1
    for i, ch in enumerate(rlabel):
        if ch == "[":
            l_sy = i
        if ch == "]":
            r_sy = i
        if ch == ">":
            n_re = len(ids)
        if r_sy > l_sy >= 0:
            part = rlabel[l_sy:r_sy + 1]
            try:
                ids.append(next(k for k, nm in enumerate(names) if nm == part))
            except StopIteration:
                raise DeckError(f"unknown particle {part} in reaction {rlabel}") from None
            l_sy = r_sy = -1
    if n_re is None:
        n_re = 0
    Eth, dE = rec.list_values(2)
    s_E, s_sig = rec.list_values(2)
    rtype = rec.list_token(20)
    kind = next((v for k, v in KINDS.items() if rtype in (k, k.lower())), 0)
    if kind == 0:
        raise DeckError(f"unknown reaction type {rtype} in {rlabel}")
    name = rtype[:4]
    while name != "----":
        name = rec.list_token(4)
    E, sig = [], []
    while True:
        line = rec.record()
        if line[:4] == "----":
            break
        rec.i -= 1
        e, s = rec.list_values(2)
        E.append(e * s_E)
        sig.append(s * s_sig)
    products = tuple(ids[n_re:])
    return Reaction(rlabel, tuple(ids[:n_re]), products, Eth * s_E, dE, kind,
                    sum(1 for p in products if p == 0), np.array(E), np.array(sig))


def _skip_surface(rec: _Records):
    name = "SURF"
    while name != "----":
        name = rec.list_token(4)
    while rec.record()[:2] != "--":
        pass


def _tracked_first(chem: Chemistry) -> Chemistry:
    order = [k for k, s in enumerate(chem.species) if s.tracked] + \
            [k for k, s in enumerate(chem.species) if not s.tracked]
    if order == list(range(len(order))):
        return chem
    new = {old: k for k, old in enumerate(order)}
    remap = lambda t: tuple(new[i] for i in t)   # noqa: E731
    reactions = [dataclasses.replace(r, reactants=remap(r.reactants), products=remap(r.products))
                 for r in chem.reactions]
    tag = lambda t: new[t] if t >= 0 else -1     # noqa: E731
    return Chemistry([chem.species[k] for k in order], [chem.density[k] for k in order],
                     reactions, tag(chem.tag_neg), tag(chem.tag_neu), tag(chem.tag_beam))


def build_tables(chem: Chemistry, ngas: float) -> dict:
    """The MCC tables of a parsed deck (ordering, mod_reactions.f90:414-558)."""
    reactions = [(r.reactants, r.products, r.Eth, r.sigma) for r in chem.reactions]
    tables = [(r.E, r.sigma) for r in chem.reactions]
    t = cfg.build_reaction_tables(chem.species, ngas, reactions, tables,
                                  neutral_density=[d for d, s in zip(chem.density, chem.species)
                                                   if not s.tracked],
                                  kinds=[(r.kind, r.n_electrons) for r in chem.reactions])
    t["names"] = [s.name for s in chem.species]
    return t


def write_reactions(species: list, fraction: list, reactions: list, path=None) -> str:
    """A deck in the reference's format: IONS and NEUTRALS sections, then one
    REACTION block per reaction (E in eV, sigma in m^2, unit scale factors).
    species[0] must be the electron; fraction is the deck's n/n0 column (the
    reader multiplies the neutrals' by ngas); reactions hold Reaction objects.
    Numbers are written with 17 significant digits, so a deck read back gives
    the same doubles."""
    def name(k):
        return "[e]" if k == 0 else f"[{species[k].name}]"
    out = []
    for title, want in (("IONS", True), ("NEUTRALS", False)):
        out += [title, "name   mass  charge  temp.(eV)  n/n0", "-" * 37]
        for k, s in enumerate(species):
            if k == 0 or s.tracked != want:
                continue
            nm = name(k)
            if len(nm) > 6:
                raise DeckError(f"species name {nm} longer than the deck's 6 characters")
            out.append(f"{nm:<6}{abs(s.mass) / cfg.AMU:.17g} {s.charge / cfg.QE:.17g} "
                       f"{s.T_eV:.17g} {fraction[k]:.17g}")
        out += ["-" * 37, ""]
    for r in reactions:
        label = "+".join(name(k) for k in r.reactants) + "->" + \
                "+".join(name(k) for k in r.products)
        out += ["REACTION", label, f" {r.Eth:.17g} {r.dE:.17g}", " 1. 1.",
                KIND_NAMES[r.kind], "-" * 60]
        out += [f" {e:.17g} {s:.17g}" for e, s in zip(r.E, r.sigma)]
        out += ["-" * 60, ""]
    text = "\n".join(out) + "\n"
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text
