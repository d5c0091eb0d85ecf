"""r-z field model: from solved potentials to the kernel's channel grids.

Restates ion_gym/physics/ionbench.py:44-96 (build_field_aware: mirror
unfold about the axis, np.gradient, one-sided differences at metal-surface
nodes, V/mm -> V/m) and the drive composition of
ion_gym/physics/build_rz.py:428-534 (static field A with DC and drive
offsets, one amplitude-baked potential per drive group, the two-point hold
"gate").  The result is the pack _fly_rec_full consumes
(build_rz.py:725-738), which ion_gym_amd.engine.Engine.upload_rz takes.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

K_SIN, K_COS, K_SQUARE = 0, 1, 2


def build_field_aware(Z_mm, U_mm, phi, ele, symmetry="cylindrical"):
    """(Z, Ue, Ez, Eu) with E in V/m on the mirror-unfolded grid
    (ionbench.py:44-96)."""
    if symmetry == "cylindrical":
        assert abs(U_mm[0]) < 1e-12
        Ue = np.concatenate([-U_mm[:0:-1], U_mm])
        pe = np.concatenate([phi[:, :0:-1], phi], axis=1)
        ee = np.concatenate([ele[:, :0:-1], ele], axis=1)
    else:
        Ue, pe, ee = U_mm.copy(), phi.copy(), ele.copy()
    dpz, dpu = np.gradient(pe, Z_mm, Ue)
    hz = Z_mm[1] - Z_mm[0]
    hu = Ue[1] - Ue[0]
    nz, nu = pe.shape
    ee = np.asarray(ee).astype(bool)
    # one-sided override at metal nodes, per axis: vacuum on exactly one side
    vac = ~ee
    vp = np.zeros_like(ee)
    vm = np.zeros_like(ee)
    vp[:-1, :] = vac[1:, :]
    vm[1:, :] = vac[:-1, :]
    fwd = ee & vp & ~vm
    bwd = ee & vm & ~vp
    dz_f = np.zeros_like(pe)
    dz_b = np.zeros_like(pe)
    dz_f[:-1, :] = (pe[1:, :] - pe[:-1, :]) / hz
    dz_b[1:, :] = (pe[1:, :] - pe[:-1, :]) / hz
    dpz = np.where(fwd, dz_f, np.where(bwd, dz_b, dpz))
    vp = np.zeros_like(ee)
    vm = np.zeros_like(ee)
    vp[:, :-1] = vac[:, 1:]
    vm[:, 1:] = vac[:, :-1]
    fwd = ee & vp & ~vm
    bwd = ee & vm & ~vp
    du_f = np.zeros_like(pe)
    du_b = np.zeros_like(pe)
    du_f[:, :-1] = (pe[:, 1:] - pe[:, :-1]) / hu
    du_b[:, 1:] = (pe[:, 1:] - pe[:, :-1]) / hu
    dpu = np.where(fwd, du_f, np.where(bwd, du_b, dpu))
    Ez = np.ascontiguousarray(-dpz * 1e3)
    Eu = np.ascontiguousarray(-dpu * 1e3)
    return (np.ascontiguousarray(Z_mm, np.float64),
            np.ascontiguousarray(Ue, np.float64), Ez, Eu)


@dataclass
class DriveGroup:
    """One drive group (the fields of RFGroupSpec the r-z route reads,
    ion_gym/io/sim_spec.py:591-692): V(t) = amplitude_v * w(t) + offset_v
    on every member electrode."""
    name: str
    waveform: str = "sin"          # sin | square | table (2-point gate)
    frequency_hz: float = 0.0
    amplitude_v: float = 0.0
    phase_deg: float = 0.0
    duty: float = 0.5
    offset_v: float = 0.0
    table_t_us: Optional[List[float]] = None
    table_v: Optional[List[float]] = None


@dataclass
class RZModel:
    """The r-z field model the engine flies (build_rz.RZModel's arrays)."""
    EzA: np.ndarray
    EuA: np.ndarray
    EzK: np.ndarray
    EuK: np.ndarray
    ch_kind: np.ndarray
    ch_om: np.ndarray
    ch_ph: np.ndarray
    ch_duty: np.ndarray
    ele: np.ndarray            # (nz, nr) labels, 0 = vacuum
    u0: float
    mm_per_gu: float
    EzG: np.ndarray = None
    EuG: np.ndarray = None
    tau_gate: float = -1.0
    extra: dict = field(default_factory=dict)

    def pack(self):
        """The kernel pack (build_rz.py:725-738): mirrored float metal
        (build_rz.py:615-616) and mirrored labels."""
        ee = np.concatenate([self.ele[:, :0:-1], self.ele], axis=1)
        return dict(EzA=self.EzA, EuA=self.EuA, EzK=self.EzK, EuK=self.EuK,
                    EzG=self.EzG, EuG=self.EuG, tau_gate=self.tau_gate,
                    ele=(ee > 0).astype(np.float64),
                    label=ee.astype(np.uint8), u0=self.u0,
                    mm=self.mm_per_gu, ch_kind=self.ch_kind,
                    ch_om=self.ch_om, ch_ph=self.ch_ph,
                    ch_duty=self.ch_duty)


def compose_terms_rz(keys, dc, members, groups):
    """The composition of an r-z model as ordered term lists, following
    build_rz.py:435-534 term by term: for each electrode idx in ``keys``
    order, fa = base_idx / 1e4 enters the static set as dc * fa, a gated
    drive's (amp * v0 + offset) * fa and a drive offset * fa, the gate set
    as amp * (v1 - v0) * fa, and each drive group's channel as amp * fa.
    Returns (terms_A, [terms per channel], terms_G, channel drives in
    order, tau_gate), each term (position of idx in ``keys``, weight);
    summing w * fa over a list in order reproduces the reference's
    accumulation exactly (the host path below and ig_fields_compose)."""
    gmap = {g.name: g for g in groups}
    tA, tG = [], []
    chan_t, chan_obj, chan_order = {}, {}, []
    tau_gate = -1.0
    for pos, idx in enumerate(keys):
        tA.append((pos, dc.get(idx, 0.0)))
        for gname in members.get(idx, ()):
            drv = gmap[gname]
            if drv.waveform == "table":
                tt, tv = list(drv.table_t_us), list(drv.table_v)
                if len(tt) != 2 or len(tv) != 2:
                    raise ValueError("r-z gate: exactly a 2-point hold table")
                if tau_gate >= 0.0 and abs(tt[1] - tau_gate) > 1e-9:
                    raise ValueError("r-z gate: two distinct gate times")
                tau_gate = tt[1]
                tA.append((pos, drv.amplitude_v * tv[0] + drv.offset_v))
                tG.append((pos, drv.amplitude_v * (tv[1] - tv[0])))
                continue
            if drv.offset_v:
                tA.append((pos, drv.offset_v))
            if drv.amplitude_v == 0.0:
                continue
            if drv.name not in chan_t:
                chan_t[drv.name] = []
                chan_obj[drv.name] = drv
                chan_order.append(drv.name)
            chan_t[drv.name].append((pos, drv.amplitude_v))
    return (tA, [chan_t[n] for n in chan_order], tG,
            [chan_obj[n] for n in chan_order], tau_gate)


def apply_terms(fas, terms, shape):
    """sum of w * fas[pos] over ``terms`` in order, from zeros."""
    acc = np.zeros(shape)
    for pos, w in terms:
        acc = acc + w * fas[pos]
    return acc


def compose_rz(bases, ele, h, dc, members, groups):
    """Compose an r-z model from per-electrode unit potentials.

    bases[idx] is electrode idx's solved potential at 1e4 V (idx 1-based,
    the reference's cache convention, build_rz.py:377-386); dc[idx] its DC
    volts; members[idx] the drive-group names it belongs to; groups the
    DriveGroup list.  Follows build_rz.py:435-534 term by term
    (compose_terms_rz)."""
    nz, nr = ele.shape
    keys = sorted(bases)
    fas = [bases[idx] / 1e4 for idx in keys]
    tA, tK, tG, drives, tau_gate = compose_terms_rz(keys, dc, members,
                                                    groups)
    A = apply_terms(fas, tA, (nz, nr))
    G = apply_terms(fas, tG, (nz, nr))
    chan_B = [apply_terms(fas, t, (nz, nr)) for t in tK]
    Z_mm = np.arange(nz) * h
    U_mm = np.arange(nr) * h
    _, Ue, EzA, EuA = build_field_aware(Z_mm, U_mm, A, ele)
    EzK_l, EuK_l = [], []
    for B in chan_B:
        _, _, a, b = build_field_aware(Z_mm, U_mm, B, ele)
        EzK_l.append(a)
        EuK_l.append(b)
    EzK = np.stack(EzK_l) if EzK_l else np.zeros((0,) + EzA.shape)
    EuK = np.stack(EuK_l) if EuK_l else np.zeros((0,) + EuA.shape)
    ch_kind = np.array([K_SQUARE if d.waveform == "square" else K_SIN
                        for d in drives], np.int64)
    ch_om = np.array([d.frequency_hz * 1e-6 * 2 * math.pi for d in drives],
                     np.float64)
    ch_ph = np.array([math.radians(d.phase_deg) for d in drives], np.float64)
    ch_duty = np.array([float(d.duty) for d in drives], np.float64)
    EzG = EuG = None
    if tau_gate >= 0.0:
        _, _, EzG, EuG = build_field_aware(Z_mm, U_mm, G, ele)
    return RZModel(EzA=EzA, EuA=EuA, EzK=EzK, EuK=EuK, ch_kind=ch_kind,
                   ch_om=ch_om, ch_ph=ch_ph, ch_duty=ch_duty, ele=ele,
                   u0=float(Ue[0]), mm_per_gu=h, EzG=EzG, EuG=EuG,
                   tau_gate=tau_gate)


def spec_terms_rz(spec):
    """(dc, members, groups) for compose_rz from a deck's geometry: each
    electrode's basis index (its ``basis``, else its 1-based position)
    with its effective DC after the ladders are resolved
    (FlightSpec.resolve_dc_groups, dc_effective) and its drive-group
    names, and the deck's drive groups (build_rz.py:435-534 reads the same
    three)."""
    spec.resolve_dc_groups()
    dc, members = {}, {}
    for pos, el in enumerate(spec.geometry.electrodes, start=1):
        idx = el.basis if el.basis is not None else pos
        dc[idx] = spec.dc_effective(el)
        members[idx] = tuple(el.group_names())
    return dc, members, list(spec.geometry.rf_groups)


def collide_labels(spec, labels):
    """The labels an r-z model carries (``ele``) from the labels a deck's
    solve fixes: an electrode declared ``is_grid`` (a mesh, a reflectron's
    entrance or back plane, a detector face) is a Dirichlet boundary of the
    solve but is transparent to ions, so its nodes get label 0 -- no splat
    on it, and the field is differenced through it (build_rz.py:397-411).
    Every other label is kept."""
    labels = np.asarray(labels)
    out = labels.copy()
    for pos, el in enumerate(spec.geometry.electrodes, start=1):
        if el.is_grid:
            out[labels == (el.basis if el.basis is not None else pos)] = 0
    return out


def compose_spec_rz(spec, bases, ele):
    """The r-z model of a deck from its solved bases (compose_rz with
    spec_terms_rz); the grid pitch is geometry.mm_per_gu.  ``ele`` is the
    deck's collide labels (collide_labels: its grids transparent)."""
    dc, members, groups = spec_terms_rz(spec)
    return compose_rz(bases, ele, spec.geometry.mm_per_gu, dc, members,
                      groups)
