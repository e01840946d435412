"""Seeded synthetic shell-model inputs in the reference's own file formats.

Writes `.sps` orbit files (read by get_orbit_info, src/binput.f90:43), OXBASH
isospin `.int` files (read by readv2bme_oxbash_iso, src/binput_tbme.f90:898;
line 1 = `nTBME eps_1 .. eps_norb` at :1087, TBME lines `a b c d J T V` at
:1171) and stdin menu scripts for the reference executable (menu, src/bmenu.f90:11;
Lanczos options, src/blanczosmain.f90:380-745).

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) with Box-Muller
Gaussians, implemented here so that every machine writes identical files.  Every
value is rounded to FP32 and printed with %.9g, so the reference's single
precision read (`real vv`, src/binput_tbme.f90:938) is exact.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; next_u64 follows the published constants."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in (0, 1): 53 random bits, offset by half a step."""
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / (1 << 53))

    def normal(self) -> float:
        """Standard normal via the Box-Muller cosine branch (one per call)."""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def fmt32(x: float) -> str:
    return "%.9g" % f32(x)


@dataclass(frozen=True)
class Orbit:
    n: int
    l: int
    j2: int  # 2j
    w: int


SD_ORBITS = (Orbit(0, 2, 5, 2), Orbit(1, 0, 1, 2), Orbit(0, 2, 3, 2))  # 0d5/2 1s1/2 0d3/2
PF_ORBITS = (Orbit(0, 3, 7, 3), Orbit(1, 1, 3, 3), Orbit(0, 3, 5, 3), Orbit(1, 1, 1, 3))
SD_SPE = (-3.95, -3.16, 1.65)
PF_SPE = (-8.6, -6.6, -2.1, -4.5)
# sd plus the 0g orbits: one parity, one W (every orbit written with w = 2,
# so no truncation applies); its Delta-m = 0 operator class has 110 members
SDG_ORBITS = SD_ORBITS + (Orbit(0, 4, 9, 2), Orbit(0, 4, 7, 2))
SDG_SPE = SD_SPE + (4.0, 6.5)
SPACES = {"sd": (SD_ORBITS, SD_SPE), "pf": (PF_ORBITS, PF_SPE), "sdg": (SDG_ORBITS, SDG_SPE)}


def sps_text(orbits: Sequence[Orbit]) -> str:
    lines = ["iso", str(len(orbits))]
    for o in orbits:
        lines.append(" %d %d %.1f %d" % (o.n, o.l, o.j2 / 2.0, o.w))
    return "\n".join(lines) + "\n"


def allowed_tbmes(orbits: Sequence[Orbit]) -> List[Tuple[int, int, int, int, int, int]]:
    """Every isospin TBME label (a,b,c,d,J,T), 1-based orbits, with a>=b, c>=d,
    pair(ab) >= pair(cd) (pair = a(a-1)/2+b, the reference's ordering,
    src/binput_tbme.f90:1291-1307), equal parity of (ab) and (cd), J within both
    couplings, and J+T odd when a==b or c==d (the legality check at :1190-1196)."""
    out = []
    norb = len(orbits)
    for a in range(1, norb + 1):
        for b in range(1, a + 1):
            pab = a * (a - 1) // 2 + b
            for c in range(1, norb + 1):
                for d in range(1, c + 1):
                    pcd = c * (c - 1) // 2 + d
                    if pcd > pab:
                        continue
                    oa, ob, oc, od = (orbits[a - 1], orbits[b - 1], orbits[c - 1], orbits[d - 1])
                    if (oa.l + ob.l + oc.l + od.l) % 2:
                        continue
                    jmin = max(abs(oa.j2 - ob.j2), abs(oc.j2 - od.j2)) // 2
                    jmax = min(oa.j2 + ob.j2, oc.j2 + od.j2) // 2
                    for J in range(jmin, jmax + 1):
                        for T in (0, 1):
                            if (a == b or c == d) and (J + T) % 2 == 0:
                                continue
                            out.append((a, b, c, d, J, T))
    return out


@dataclass
class Interaction:
    orbits: Tuple[Orbit, ...]
    spe: Tuple[float, ...]
    tbmes: List[Tuple[int, int, int, int, int, int, float]] = field(default_factory=list)

    def int_text(self) -> str:
        head = "%d " % len(self.tbmes) + " ".join(fmt32(e) for e in self.spe)
        lines = ["! seeded synthetic isospin interaction", head]
        for (a, b, c, d, J, T, v) in self.tbmes:
            lines.append("%d %d %d %d %d %d %s" % (a, b, c, d, J, T, fmt32(v)))
        return "\n".join(lines) + "\n"


def gaussian_interaction(orbits, spe, seed: int, spe_sigma: float = 0.0) -> Interaction:
    """TBMEs ~ N(0,1) MeV off-diagonal, N(0,2) (sigma sqrt 2) when (ab)==(cd);
    SPEs fixed (spe_sigma == 0) or drawn N(0, spe_sigma) first."""
    rng = SplitMix64(seed)
    if spe_sigma > 0.0:
        spe = tuple(f32(spe_sigma * rng.normal()) for _ in orbits)
    else:
        spe = tuple(f32(e) for e in spe)
    tb = []
    for (a, b, c, d, J, T) in allowed_tbmes(orbits):
        sigma = math.sqrt(2.0) if (a, b) == (c, d) else 1.0
        tb.append((a, b, c, d, J, T, f32(sigma * rng.normal())))
    return Interaction(tuple(orbits), spe, tb)


@dataclass
class Case:
    """One solve: model space, nucleons, interaction and Lanczos settings."""
    name: str
    space: str          # "sd", "pf" or "sdg"
    Z: int
    N: int
    jz2: int
    interaction: Interaction
    nkeep: int
    niter: int
    thick: int = 0      # 0 = plain Lanczos; else number of thick-restart vectors kept
    restarts: int = 20  # maximum number of thick restarts

    @property
    def orbits(self):
        return self.interaction.orbits


def config_case(cfg: int, member: int = 0) -> Case:
    """The five benchmark configurations (sd 4p4n, sd 6p6n, sd 4p4n ensemble,
    pf 4p4n, pf 6p6n), all at M=0 with seeded Gaussian two-body ensembles."""
    if cfg in (0, 1):
        it = gaussian_interaction(SD_ORBITS, SD_SPE, seed=1000)
        zn = 4 if cfg == 0 else 6
        return Case("sd%dp%dn" % (zn, zn), "sd", zn, zn, 0, it,
                    nkeep=5 if cfg == 0 else 10, niter=150)
    if cfg == 2:
        it = gaussian_interaction(SD_ORBITS, SD_SPE, seed=2000 + member, spe_sigma=2.0)
        return Case("sdens%02d" % member, "sd", 4, 4, 0, it, nkeep=5, niter=150)
    if cfg in (3, 4):
        it = gaussian_interaction(PF_ORBITS, PF_SPE, seed=3000)
        if cfg == 3:
            return Case("pf4p4n", "pf", 4, 4, 0, it, nkeep=10, niter=60, thick=15)
        return Case("pf6p6n", "pf", 6, 6, 0, it, nkeep=5, niter=40, thick=10)
    raise ValueError("unknown configuration %r" % cfg)


def small_case(space: str, Z: int, N: int, jz2: int = 0, seed: int = 77,
               nkeep: int = 5, niter: int = 60) -> Case:
    orbits, spe = SPACES[space]
    it = gaussian_interaction(orbits, spe, seed=seed)
    return Case("%s%dp%dn" % (space, Z, N), space, Z, N, jz2, it, nkeep=nkeep, niter=niter)


def write_case(case: Case, directory: str) -> Tuple[str, str]:
    """Write <space>.sps and <name>.int into `directory`; return their stems."""
    os.makedirs(directory, exist_ok=True)
    sps = case.space
    with open(os.path.join(directory, sps + ".sps"), "w") as f:
        f.write(sps_text(case.orbits))
    with open(os.path.join(directory, case.name + ".int"), "w") as f:
        f.write(case.interaction.int_text())
    return sps, case.name


def menu_script(case: Case, option: str = "n", lanczos: str = None, block=None) -> str:
    """stdin answers for the reference executable: menu option, output name,
    .sps, Z N, 2Jz, interaction, scaling `1 1 1 1` (no mass scaling,
    src/binput_tbme.f90:1027-1031), `end`, then the diagonalization option;
    block = (dimblock, block iterations) for the block options bf / bd
    (src/blanczosmain.f90:822-872)."""
    sps, intf = case.space, case.name
    lines = [option, case.name, sps, "%d %d" % (case.Z, case.N), str(case.jz2),
             intf, "1 1 1 1", "end"]
    if option == "wh":
        return "\n".join(lines) + "\n"
    if lanczos is None:
        lanczos = "tf" if case.thick else "lf"
    lines.append(lanczos)
    if lanczos.startswith("b"):
        lines.append("%d %d %d" % (case.nkeep, block[0], block[1]))
        return "\n".join(lines) + "\n"
    lines.append("%d %d" % (case.nkeep, case.niter))
    if lanczos.startswith("t"):
        lines.append(str(case.thick))
        lines.append(str(case.restarts))
    return "\n".join(lines) + "\n"
