"""Per-event weights per resonance group: the counterpart of
RooDalitzAmplitude::storeResonanceWeights (src/RooDalitzAmplitude.cxx:985-1292).

For every group the helicity couplings of the resonances outside it are set
to zero (the include_resonance switch of ConfigureCouplings,
src/RooDalitzAmplitude.cxx:2707-2733, :2789-2814, :2951) and the per-event
|A|^2 is evaluated as in the fit (pwa_eval: smeared data in the m(J/psi p)
windows, MC weighted), giving the reference's data branches (amp2 and one
per group) and MC branches (amp2_weight and one per group).  Groups, in the
reference's naming:

  amp2                     every resonance
  allLz, allPc, allZcs     one chain
  LZ_JP_<2J><p|m>          Lambda* of one J^P;  LZ_high_J: Lambda* with 2J > 3
  Pc_JP_<2J><p|m>          Pc of one J^P;  allLz_Pc_JP_<..>: with every Lambda*
  Zcs_JP_<2J><p|m>         Zcs of one J^P
  Ls_<i>, Pc_<i>, Zcs_<i>  one Breit-Wigner / (revised) K-matrix resonance with
                           m0 >= 0 (i: index within its chain; the reference
                           takes the name from the RooArgList, here `names`)
  NR_Ls_<i>, NR_Pc_<i>     one non-resonant term
"""
import numpy as np

from . import abi


def _chain_members(desc):
    nl, npc, nx = desc.n_lambda, desc.n_pc, desc.n_zcs
    return list(range(nl)), list(range(nl, nl + npc)), list(range(nl + npc, nl + npc + nx))


def _jp_groups(desc, members):
    seen = []
    for r in members:
        jp = (desc.spin2[r], desc.parity[r])
        if jp not in seen:
            seen.append(jp)
    return [(s2, p, [r for r in members if (desc.spin2[r], desc.parity[r]) == (s2, p)]) for s2, p in seen]


def resonance_groups(desc, params, names=None):
    """Ordered {group name: resonance indices} as storeResonanceWeights builds
    them; `names` optionally maps a resonance index to its name suffix."""
    lam, pc, zcs = _chain_members(desc)
    names = names or {}
    g = {"amp2": lam + pc + zcs, "allLz": list(lam), "allPc": list(pc), "allZcs": list(zcs)}
    for s2, p, rs in _jp_groups(desc, lam):
        g[f"LZ_JP_{s2}{'p' if p > 0 else 'm'}"] = rs
    g["LZ_high_J"] = [r for r in lam if desc.spin2[r] > 3]
    for s2, p, rs in _jp_groups(desc, pc):
        key = f"Pc_JP_{s2}{'p' if p > 0 else 'm'}"
        g[key] = rs
        g["allLz_" + key] = list(lam) + rs
    for s2, p, rs in _jp_groups(desc, zcs):
        g[f"Zcs_JP_{s2}{'p' if p > 0 else 'm'}"] = rs
    single = (abi.PWA_BREITWIGNER, abi.PWA_KMATRIX, abi.PWA_REVISEDKMATRIX)
    for tag, members, base in (("Ls", lam, 0), ("Pc", pc, len(lam)), ("Zcs", zcs, len(lam) + len(pc))):
        for r in members:
            m0 = params[desc.param_offset(r) + desc.param_size(r) - 5]
            if desc.type[r] in single and m0 >= 0:
                g[f"{tag}_{names.get(r, r - base)}"] = [r]
            elif desc.type[r] == abi.PWA_NONRESONANT:
                g[f"NR_{tag}_{names.get(r, r - base)}"] = [r]
    return g


def exclude(desc, params, keep):
    """Parameters with every resonance outside `keep` switched off: its
    helicity couplings become zero (Lambda*: the overall scale; Pc and Zcs:
    their LS couplings)."""
    p = np.array(params, dtype=np.float64)
    keep = set(keep)
    for r in range(desc.n_res):
        if r in keep:
            continue
        o, size = desc.param_offset(r), desc.param_size(r)
        ch = desc.chain(r)
        if ch == abi.PWA_CHAIN_LAMBDA:
            p[o + size - 7] = p[o + size - 6] = 0.0
        elif ch == abi.PWA_CHAIN_PC:
            p[o + 2:o + 8] = 0.0
        else:
            p[o:o + 8] = 0.0
    return p


def resonance_weights(ctx, desc, params, groups=None, names=None):
    """{group: (per-data-event |A|^2, per-MC-event w |A|^2)} in caller order."""
    groups = groups if groups is not None else resonance_groups(desc, params, names)
    out = {}
    for name, keep in groups.items():
        _, _, fxd, fxm = ctx.eval(exclude(desc, params, keep), want_fx=True)
        out[name] = (fxd, fxm)
    return out
