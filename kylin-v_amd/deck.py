"""Reference input decks (kln-holstein's `-i` file) read into the engine's MPO builder.

read_raw follows holstein/input.h:85-155: a section starts at a line that contains
"electron", "vibration", "1st-order", "2nd-order" or "reorder", and its numbers follow as
whitespace-separated tokens.  Electronic and vibrational labels are 1-based in the file;
internally an electronic state e is label e - 1, vibration v is label v + NumEle - 1 and its
auxiliary (purification) site is the negated label, placed right after it by the reordering
(input.h:140-150).

lattice_from_deck builds the operator that holstein/lattice.h:183-339 (totals) assembles by
MPO products, as a list of product terms for model.mpo_blocks:
  H_e = sum val |e1><e2|                                   (no Jordan-Wigner sign)
  H_v = sum w b+ b                                         (no zero-point term)
  H_1 = sum val sqrt(0.5) |e1><e2| (b+ a_aux + b a+_aux)
  H_2 = sum val 0.5 |e1><e2| (b1+ a1 + b1 a1+)(b2+ a2 + b2 a2+)
with b+ carrying sqrt(i+1) and the auxiliary ladder operators carrying 1
(holstein/local.h:43-86), exciton sites {0, 1}, vibrational and auxiliary sites
{0..occ} and the auxiliary sites starting full (holstein/lattice.h:83-124).  The result is
the projected-purification layout the reference runs.
"""
from __future__ import annotations

import numpy as np

from . import model as KM


def _tokens(lines):
    for line in lines:
        for tok in line.split():
            yield tok


def read_raw(path_or_text):
    """Parse a deck (a path, or the deck text itself when it contains a newline)."""
    if "\n" in path_or_text:
        text = path_or_text
    else:
        with open(path_or_text) as f:
            text = f.read()
    lines = text.splitlines()
    cfg = dict(num_ele=0, num_vib=0, exc=[], freq={}, dim={}, linear=[], bilinear=[], site_to_lab=[])
    i = 0
    while i < len(lines):
        head = lines[i]
        i += 1
        if not any(k in head for k in ("electron", "vibration", "1st-order", "2nd-order", "reorder")):
            continue
        # the numbers of this section run until the next section header
        j = i
        while j < len(lines) and not any(k in lines[j] for k in
                                         ("electron", "vibration", "1st-order", "2nd-order", "reorder")):
            j += 1
        tok = _tokens(lines[i:j])
        i = j
        if "electron" in head:
            cfg["num_ele"], n = int(next(tok)), int(next(tok))
            for _ in range(n):
                e1, e2, val = int(next(tok)) - 1, int(next(tok)) - 1, float(next(tok))
                cfg["exc"].append((e1, e2, val))
                cfg["dim"][e1] = 2
        elif "vibration" in head:
            cfg["num_vib"] = int(next(tok))
            for _ in range(cfg["num_vib"]):
                v, fq, occ = int(next(tok)), float(next(tok)), int(next(tok))
                lab = v + cfg["num_ele"] - 1
                cfg["freq"][lab] = fq
                cfg["dim"][lab] = occ + 1
                cfg["dim"][-lab] = occ + 1
        elif "1st-order" in head:
            for _ in range(int(next(tok))):
                v, e1, e2, val = int(next(tok)), int(next(tok)), int(next(tok)), float(next(tok))
                cfg["linear"].append((v + cfg["num_ele"] - 1, e1 - 1, e2 - 1, val))
        elif "2nd-order" in head:
            for _ in range(int(next(tok))):
                v1, v2 = int(next(tok)), int(next(tok))
                e1, e2, val = int(next(tok)), int(next(tok)), float(next(tok))
                cfg["bilinear"].append((v1 + cfg["num_ele"] - 1, v2 + cfg["num_ele"] - 1, e1 - 1, e2 - 1, val))
        else:  # reorder
            for _ in range(cfg["num_ele"] + cfg["num_vib"]):
                lab = int(next(tok))
                cfg["site_to_lab"].append(lab)
                if lab >= cfg["num_ele"]:
                    cfg["site_to_lab"].append(-lab)
    if not cfg["site_to_lab"]:
        raise ValueError("deck has no reordering part")
    return cfg


def _pp_sites(cfg):
    """Site objects in the deck's MPS order and the label -> site index map."""
    e = KM._exciton()
    sites, lab_to_site = [], {}
    cache = {}
    for i, lab in enumerate(cfg["site_to_lab"]):
        lab_to_site[lab] = i
        if 0 <= lab < cfg["num_ele"]:
            sites.append(e)
            continue
        L = cfg["dim"][lab]
        key = ("phys" if lab > 0 else "aux", L)
        if key not in cache:
            b, bd = KM._boson(L)
            up = np.diag(np.ones(L - 1), -1)
            cache[key] = (KM.Site(list(range(L)), {"I": np.eye(L), "bdag": bd, "b": b, "nb": bd @ b})
                          if lab > 0 else
                          KM.Site(list(range(L)), {"I": np.eye(L), "araise": up, "alower": up.T}))
        sites.append(cache[key])
    return sites, lab_to_site


def lattice_from_deck(cfg, exciton=0):
    """(sites, terms, init, exciton_sites) for model.mpo_blocks / product_state_blocks, with the
    exciton in electronic state `exciton` (0-based; the reference's State/Ele-<exciton+1>)."""
    sites, l2s = _pp_sites(cfg)
    terms = []

    def ele(e1, e2):
        s1, s2 = l2s[e1], l2s[e2]
        return [(s1, "n")] if s1 == s2 else [(s1, "cdag"), (s2, "c")]

    def term(coef, ops):  # model.mpo_blocks takes each term's operators in site order
        return (coef, sorted(ops))

    def ladder(v):
        p, a = l2s[v], l2s[-v]
        return [[(p, "bdag"), (a, "alower")], [(p, "b"), (a, "araise")]]

    for e1, e2, val in cfg["exc"]:
        terms.append(term(val, ele(e1, e2)))
    for v in sorted(cfg["freq"]):
        terms.append(term(cfg["freq"][v], [(l2s[v], "nb")]))
    for v, e1, e2, val in cfg["linear"]:
        for lad in ladder(v):
            terms.append(term(val * np.sqrt(0.5), ele(e1, e2) + lad))
    for v1, v2, e1, e2, val in cfg["bilinear"]:
        if v1 == v2:
            raise ValueError("2nd-order terms on one mode (v1 == v2) are not supported")
        for l1 in ladder(v1):
            for l2 in ladder(v2):
                terms.append(term(val * 0.5, ele(e1, e2) + l1 + l2))
    init = []
    for lab in cfg["site_to_lab"]:
        if 0 <= lab < cfg["num_ele"]:
            init.append(1 if lab == exciton else 0)
        elif lab > 0:
            init.append(0)
        else:
            init.append(cfg["dim"][lab] - 1)  # the auxiliary site starts full
    ex = [l2s[e] for e in range(cfg["num_ele"])]
    return sites, terms, init, ex
