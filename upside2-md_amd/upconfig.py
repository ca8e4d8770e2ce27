"""Writer for ff_2.1 `.up` configuration files (the `/input` layout).

This is the data format on the input side of the force engine: the reference
builds it with py/upside_config.py (PyTables, absent here).  This module builds
the same node groups from numpy arrays and writes them with h5lite.  Each
writer cites the reference function whose layout it reproduces.

Node set (ff_2.1, `--dynamic-rotamer-1body`, fixed rotamer placement, sigmoid
environment coupling, backbone environment), as in
example/15.SpatialTransformation/inputs/chig.up:
  springs, rama_map_pot(+_ref), backbone_pairs, infer_H_O, protein_hbond,
  hbond_energy, rotamer placement, hbond_coverage(+_hydrophobe), rotamer,
  environment coverage + sigmoid coupling, backbone environment, weighted_pos,
  Dihedral_phi/psi, rama_coord, affine_alignment.

The Rama library of ff_2.1 (parameters/common/rama.dat) is not part of the
reference tree, so `standin_rama_maps` makes a seeded smooth stand-in with the
normalisation of write_rama_map_pot (py/upside_config.py:761-846).  A caller
that has real maps (e.g. chig.up's) passes them as `rama_pot`.
"""
import os

import numpy as np

try:
    from . import h5lite
except ImportError:  # loaded as a top-level module
    import h5lite

_HERE = os.path.dirname(os.path.abspath(__file__))
PARAM_PACK = os.path.join(_HERE, "data", "ff21_params.npz")

three_letter_aa = dict(A='ALA', C='CYS', D='ASP', E='GLU', F='PHE', G='GLY', H='HIS', I='ILE',
                       K='LYS', L='LEU', M='MET', N='ASN', P='PRO', Q='GLN', R='ARG', S='SER',
                       T='THR', V='VAL', W='TRP', Y='TYR')
RESTYPES = sorted(three_letter_aa.values())
deg = np.pi / 180.0
n_bit_rotamer = 4

# py/upside_config.py:56-60 (backbone_group4 is the grouping write_rotamer_backbone uses)
_backbone_group4 = dict((r, 0) for r in RESTYPES)
_backbone_group4.update(GLY=1, PRO=2)

# backbone reference geometry N, CA, C, CB (py/upside_config.py:320-323,1005-1008)
_REF_NCAC_CB = np.array([(-1.19280531, -0.83127186, 0.),
                         (0., 0., 0.),
                         (1.25222632, -0.87268266, 0.),
                         (0., 0.94375626, 1.2068012)])


def load_params(path=PARAM_PACK):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ------------------------------------------------------------------ geometry
def _tab_matrices(phi, theta, bond_length):
    """Torsion-angle-bond affine matrices (py/upside_config.py:140-158)."""
    cp, sp = np.cos(phi), np.sin(phi)
    ct, st = np.cos(theta), np.sin(theta)
    l = bond_length
    r = np.zeros(phi.shape + (4, 4))
    r[..., 0, :] = np.stack([-ct, -st, 0 * ct, -l * ct], -1)
    r[..., 1, :] = np.stack([cp * st, -cp * ct, -sp, l * cp * st], -1)
    r[..., 2, :] = np.stack([sp * st, -sp * ct, cp, l * sp * st], -1)
    r[..., 3, 3] = 1.0
    return r


def random_initial_config(n_res, rng):
    """Backbone from uniformly random (phi, psi), trans omega and ideal angles and
    bonds (py/upside_config.py:160-201), with a seeded generator."""
    rama = rng.random((n_res, 3)) * 2 * np.pi - np.pi
    rama[:, 2] = np.pi
    angles = np.zeros_like(rama)
    lengths = np.zeros_like(rama)
    angles[:, 0], angles[:, 1], angles[:, 2] = 120.0 * deg, 120.0 * deg, 109.5 * deg
    lengths[:, 0], lengths[:, 1], lengths[:, 2] = 1.453, 1.526, 1.300
    n_atom = 3 * n_res
    t = np.zeros(n_atom)
    t[3::3] = rama[:-1, 1]
    t[4::3] = rama[:-1, 2]
    t[5::3] = rama[1:, 0]
    transforms = _tab_matrices(t, angles.ravel(), lengths.ravel())
    cur = np.eye(4)
    pos = np.zeros((n_atom, 3))
    for i, m in enumerate(transforms):
        cur = cur.dot(m)
        pos[i] = cur[:3, 3]
    return pos


def random_sequence(n_res, rng):
    return [RESTYPES[i] for i in rng.integers(0, 20, size=n_res)]


# ----------------------------------------------------------- Rama stand-in
def standin_rama_maps(seq, seed=0, n_grid=72):
    """Deterministic smooth periodic Rama potentials, one 72x72 map per residue.

    Each residue type gets a mixture of periodic (von Mises) basins near the
    alpha, beta, PPII and left-handed regions, with weights drawn from a
    generator seeded by (seed, residue type).  Normalisation follows
    read_rama_maps_and_weights (py/upside_config.py:743-745):
    sum(exp(-pot)) = 1 over the grid.
    """
    phi = np.linspace(-np.pi, np.pi, n_grid, endpoint=False)[:, None]
    psi = np.linspace(-np.pi, np.pi, n_grid, endpoint=False)[None, :]
    centers = np.array([(-63., -43.), (-120., 130.), (-75., 145.), (57., 47.)]) * deg
    cache = {}
    out = np.zeros((len(seq), n_grid, n_grid))
    for i, r in enumerate(seq):
        if r not in cache:
            g = np.random.default_rng([seed, RESTYPES.index(r if r != 'CPR' else 'PRO')])
            w = g.dirichlet(np.ones(4)) + 0.05
            if r != 'GLY':
                w[3] *= 0.1
            kappa = 2.0 + 3.0 * g.random(4)
            p = np.zeros((n_grid, n_grid))
            for (c0, c1), wk, kk in zip(centers, w, kappa):
                p += wk * np.exp(kk * (np.cos(phi - c0) - 1) + kk * (np.cos(psi - c1) - 1))
            p += 1e-3
            pot = -np.log(p)
            pot -= -np.log(np.exp(-pot).sum())
            cache[r] = pot
        out[i] = cache[r]
    return out


# ------------------------------------------------------------------ writers
class Graph:
    """In-memory `/input` tree: node groups with datasets and attributes."""

    def __init__(self):
        self.groups = {}   # path -> {attr: value}
        self.dsets = {}    # path -> array
        self.order = []

    def group(self, name, arguments=None, **attrs):
        path = "/input/potential/" + name
        a = dict(attrs)
        if arguments is not None:
            a["arguments"] = list(arguments)
        self.groups[path] = a
        self.order.append(path)
        return path

    def subgroup(self, path, **attrs):
        self.groups[path] = dict(attrs)
        self.order.append(path)
        return path

    def ds(self, path, arr):
        self.dsets[path] = np.asarray(arr)


def _spring_nodes(g, seq, n_atom, bond_k=48., angle_k=175., omega_k=30.):
    """write_dist_spring / write_angle_spring / write_omega_spring1
    (py/upside_config.py:453-569)."""
    i = np.arange(n_atom - 1)
    p = g.group("Distance3D", ["pos", "pos"])
    g.ds(p + "/id", np.column_stack((i, i + 1)))
    eq = np.zeros(n_atom - 1)
    eq[0::3], eq[1::3], eq[2::3] = 1.453, 1.526, 1.300
    p = g.group("Spring_bond", ["Distance3D"], integrator_level=0, dim1=0, pbc=0, box_len=0)
    g.ds(p + "/id", np.arange(n_atom - 1))
    g.ds(p + "/equil_dist", eq)
    g.ds(p + "/spring_const", bond_k * np.ones(n_atom - 1))

    i = np.arange(n_atom - 2)
    p = g.group("Angle", ["pos"])
    g.ds(p + "/id", np.column_stack((i, i + 2, i + 1)))
    eq = np.zeros(n_atom - 2)
    eq[0::3], eq[1::3], eq[2::3] = np.cos(109.5 * deg), np.cos(120.0 * deg), np.cos(120.0 * deg)
    p = g.group("Spring_angle", ["Angle"], integrator_level=0, dim1=0, pbc=0, box_len=0)
    g.ds(p + "/id", np.arange(n_atom - 2))
    g.ds(p + "/equil_dist", eq)
    g.ds(p + "/spring_const", angle_k * np.ones(n_atom - 2))

    i = np.arange(1, n_atom - 3, 3)
    p = g.group("Dihedral_omega", ["pos"])
    g.ds(p + "/id", np.column_stack((i, i + 1, i + 2, i + 3)))
    seq = np.array(seq)
    target = np.where(seq[1:] == 'CPR', 0. * deg, 180. * deg)
    p = g.group("Spring_omega", ["Dihedral_omega"], integrator_level=0, dim1=0, pbc=1, box_len=np.pi * 2)
    g.ds(p + "/id", np.arange(len(i)))
    g.ds(p + "/equil_dist", target)
    g.ds(p + "/spring_const", omega_k * np.ones(len(i)))


def _rama_nodes(g, seq, rama_pot, params, normalize=True):
    """write_rama_map_pot + rama_map_pot_ref (py/upside_config.py:761-846,2401-2413)."""
    rama_pot = np.array(rama_pot, dtype=np.float64)
    if normalize:
        rama_pot -= (rama_pot * np.exp(-rama_pot)).sum(axis=(-2, -1), keepdims=True)
    p = g.group("rama_map_pot", ["rama_coord"], integrator_level=0, sheet_eps=1e-2)
    n = len(seq)
    g.ds(p + "/residue_id", np.arange(n))
    g.ds(p + "/rama_map_id", np.arange(rama_pot.shape[0]))
    g.ds(p + "/rama_map_id_all", np.arange(rama_pot.shape[0]))
    g.ds(p + "/rama_pot", rama_pot)

    ref = np.log(params["rama_reference"])
    ref = ref - ref.mean()
    p = g.group("rama_map_pot_ref", ["rama_coord"], log_pot=0)
    g.ds(p + "/residue_id", np.arange(n))
    g.ds(p + "/rama_map_id", np.zeros(n, dtype=np.int32))
    g.ds(p + "/rama_pot", ref[None])


def _backbone_pairs(g, seq):
    """write_backbone_pair (py/upside_config.py:998-1016)."""
    n = len(seq)
    ref = np.zeros((n, 4, 3))
    ref[:] = _REF_NCAC_CB
    ref[np.array(seq) == 'GLY', 3] = np.nan
    ref -= ref[:, :3].mean(axis=1)[:, None]
    p = g.group("backbone_pairs", ["affine_alignment"], integrator_level=1)
    g.ds(p + "/id", np.arange(n))
    g.ds(p + "/ref_pos", ref)
    g.ds(p + "/n_atom", np.isfinite(ref.sum(axis=-1)).sum(axis=-1))


def _hbond_nodes(g, seq, params):
    """write_infer_H_O, write_count_hbond, write_short_hbond
    (py/upside_config.py:332-368,1018-1075)."""
    n = len(seq)
    donors = np.array([i for i in range(n) if i > 0 and seq[i] != 'PRO'])
    acceptors = np.array([i for i in range(n) if i < n - 1])
    p = g.group("infer_H_O", ["pos"])
    d = g.subgroup(p + "/donors")
    a = g.subgroup(p + "/acceptors")
    g.ds(d + "/residue", donors)
    g.ds(a + "/residue", acceptors)
    g.ds(d + "/bond_length", 0.88 * np.ones(len(donors)))
    g.ds(a + "/bond_length", 1.24 * np.ones(len(acceptors)))
    g.ds(d + "/id", np.array((-1, 0, 1))[None, :] + 3 * donors[:, None])
    g.ds(a + "/id", np.array((1, 2, 3))[None, :] + 3 * acceptors[:, None])

    nd, na = len(donors), len(acceptors)
    p = g.group("protein_hbond", ["infer_H_O"])
    g.ds(p + "/index1", np.arange(nd))
    g.ds(p + "/type1", np.zeros(nd, dtype=np.int32))
    g.ds(p + "/id1", donors)
    g.ds(p + "/index2", np.arange(nd, nd + na))
    g.ds(p + "/type2", np.zeros(na, dtype=np.int32))
    g.ds(p + "/id2", acceptors)
    g.ds(p + "/interaction_param", np.array([[[1.4, 1. / 0.10, 2.5, 1. / 0.125, 0.682, 1. / 0.05, 0., 0.]]]))

    p = g.group("hbond_energy", ["protein_hbond", "rama_coord"], integrator_level=1)
    g.ds(p + "/parameters", params["hbond/parameter"])
    g.ds(p + "/donor_resid", donors)
    g.ds(p + "/acceptor_resid", acceptors)
    g.ds(p + "/rama_resid", np.arange(n))
    return donors, acceptors


def _rotamer_placement(g, seq, params):
    """write_rotamer_placement with fixed placement and dynamic 1-body energies
    (py/upside_config.py:1077-1202)."""
    restype_num = dict((x.decode(), i) for i, x in enumerate(params["sidechain/restype_order"]))
    start_stop = params["sidechain/rotamer_start_stop_bead"]
    placement_pos = params["sidechain/rotamer_center_fixed"]
    placement_energy = -np.log(params["sidechain/rotamer_prob"].transpose((2, 0, 1)))[..., None]
    rama_residue, affine_residue, layer_index, beadtype_seq, id_seq = [], [], [], [], []
    count_by_n_rot = {}
    for rnum, aa in enumerate(seq):
        aa = 'PRO' if aa == 'CPR' else aa
        start, stop, n_bead = start_stop[restype_num[aa]]
        n_rot = (stop - start) // n_bead
        count_by_n_rot.setdefault(n_rot, 0)
        base_id = (count_by_n_rot[n_rot] << n_bit_rotamer) + n_rot
        count_by_n_rot[n_rot] += 1
        rama_residue.extend([rnum] * (stop - start))
        affine_residue.extend([rnum] * (stop - start))
        layer_index.extend(np.arange(start, stop))
        beadtype_seq.extend(['%s_%i' % (aa, i) for i in range(n_bead)] * n_rot)
        id_seq.extend(np.arange(stop - start) // n_bead + (base_id << n_bit_rotamer))

    sc = "placement_fixed_point_vector_only"
    p = g.group(sc, ["affine_alignment"])
    g.ds(p + "/rama_residue", np.array(rama_residue))
    g.ds(p + "/affine_residue", np.array(affine_residue))
    g.ds(p + "/layer_index", np.array(layer_index))
    g.ds(p + "/placement_data", placement_pos[..., :6])
    g.ds(p + "/beadtype_seq", np.array(beadtype_seq, dtype="S"))
    g.ds(p + "/id_seq", np.array(id_seq))
    g.ds(p + "/fix_rotamer", np.zeros((0,), dtype=np.int64))

    pl = "placement_scalar"
    p = g.group(pl, ["affine_alignment", "rama_coord"])
    g.ds(p + "/rama_residue", np.array(rama_residue))
    g.ds(p + "/affine_residue", np.array(affine_residue))
    g.ds(p + "/layer_index", np.array(layer_index))
    g.ds(p + "/placement_data", placement_energy)
    return sc, pl, np.array(beadtype_seq), np.array(affine_residue), np.array(id_seq)


def _rotamer_backbone(g, seq, params, donors, acceptors, sc, beadtypes, sc_resnum):
    """write_rotamer_backbone (py/upside_config.py:1204-1299)."""
    n = len(seq)
    num_group = len(set(_backbone_group4.values()))
    bead_num = dict((k.decode(), i) for i, k in enumerate(params["sidechain/bead_order"]))
    sc_type = np.array([bead_num[s] for s in beadtypes])
    sc_index = np.arange(len(beadtypes))

    def bbt_of(resid):
        bbt = _backbone_group4[seq[resid] if seq[resid] != 'CPR' else 'PRO']
        if resid < n - 1 and seq[resid + 1] in ('PRO', 'CPR'):
            bbt = num_group
        return bbt

    hb_type = [bbt_of(r) * 2 + 0 for r in donors] + [bbt_of(r) * 2 + 1 for r in acceptors]
    p = g.group("hbond_coverage", ["protein_hbond", sc])
    g.ds(p + "/interaction_param", params["sidechain/coverage_interaction"])
    g.ds(p + "/index1", np.arange(len(donors) + len(acceptors)))
    g.ds(p + "/type1", np.array(hb_type))
    g.ds(p + "/id1", np.concatenate([donors, acceptors]))
    g.ds(p + "/index2", sc_index)
    g.ds(p + "/type2", sc_type)
    g.ds(p + "/id2", sc_resnum)

    bb_index = np.arange(3 * n)
    bb_resnum = bb_index / 3  # float, as the reference writes it; the loader truncates
    p = g.group("placement_fixed_point_vector_scalar", ["affine_alignment"])
    g.ds(p + "/affine_residue", bb_resnum)
    g.ds(p + "/layer_index", bb_index % 3)
    g.ds(p + "/placement_data", params["sidechain/hydrophobe_placement"])

    bb_type = [bbt_of(r) * 3 + i for r in range(n) for i in (0, 1, 2)]
    p = g.group("hbond_coverage_hydrophobe", ["placement_fixed_point_vector_scalar", sc])
    g.ds(p + "/interaction_param", params["sidechain/hydrophobe_interaction"])
    g.ds(p + "/index1", bb_index)
    g.ds(p + "/type1", np.array(bb_type))
    g.ds(p + "/id1", bb_resnum)
    g.ds(p + "/index2", sc_index)
    g.ds(p + "/type2", sc_type)
    g.ds(p + "/id2", sc_resnum)


def _rotamer(g, params, sc, pl, beadtypes, id_seq, damping=0.4, max_iter=1000, tol=1e-3, chunk=2):
    """write_rotamer (py/upside_config.py:1302-1335)."""
    args = [sc, pl, "hbond_coverage", "hbond_coverage_hydrophobe"]
    p = g.group("rotamer", args, integrator_level=1, max_iter=max_iter, tol=tol, damping=damping,
                iteration_chunk_size=chunk)
    pg = g.subgroup(p + "/pair_interaction")
    bead_num = dict((k.decode(), i) for i, k in enumerate(params["sidechain/bead_order"]))
    g.ds(pg + "/interaction_param", params["sidechain/pair_interaction"])
    g.ds(pg + "/index", np.arange(len(beadtypes)))
    g.ds(pg + "/type", np.array([bead_num[s] for s in beadtypes]))
    g.ds(pg + "/id", id_seq)


def _environment(g, seq, params, sc, sc_resnum):
    """write_environment with sigmoid coupling (py/upside_config.py:1341-1426)."""
    n = len(seq)
    restype_order = dict((x.decode(), i) for i, x in enumerate(params["environment/restype_order"]))
    ref = _REF_NCAC_CB - _REF_NCAC_CB[:3].mean(axis=0, keepdims=True)
    pd = np.zeros((1, 6))
    pd[0, 0:3] = ref[3]
    v = ref[3] - ref[1]
    pd[0, 3:6] = v / np.sqrt((v ** 2).sum())
    p = g.group("placement_fixed_point_vector_only_CB", ["affine_alignment"])
    g.ds(p + "/affine_residue", np.arange(n))
    g.ds(p + "/layer_index", np.zeros(n, dtype=np.int32))
    g.ds(p + "/placement_data", pd)

    aa = np.array([restype_order['PRO' if s == 'CPR' else s] for s in seq])
    n_sc = len(sc_resnum)
    p = g.group("environment_coverage_sc", ["placement_fixed_point_vector_only_CB", "weighted_pos"],
                num_aa_types=20)
    g.ds(p + "/index1", np.arange(n))
    g.ds(p + "/type1", aa)
    g.ds(p + "/id1", np.arange(n))
    g.ds(p + "/index2", np.arange(n_sc))
    g.ds(p + "/type2", 0 * np.arange(n_sc))
    g.ds(p + "/id2", sc_resnum)
    g.ds(p + "/interaction_param", params["environment/coverage_param"])
    g.ds(p + "/aa_types", aa)

    p = g.group("sigmoid_coupling_environment", ["environment_coverage_sc"], integrator_level=1,
                number_independent_weights=20)
    g.ds(p + "/coupling_types", aa)
    g.ds(p + "/weights", params["environment/weights"])
    g.ds(p + "/scale", params["environment/scale"])
    g.ds(p + "/center", params["environment/center"])
    g.ds(p + "/sharpness", params["environment/sharpness"])


def _bb_environment(g, seq, params, donors, acceptors, sc_resnum):
    """write_bb_environment without heavy-atom coverage (py/upside_config.py:1428-1541)."""
    restype_order = dict((x.decode(), i) for i, x in enumerate(params["environment/restype_order"]))
    res_id = np.concatenate([donors, acceptors])
    nd, na = len(donors), len(acceptors)
    n_sc = len(sc_resnum)
    p = g.group("environment_coverage_hb", ["infer_H_O", "weighted_pos"], num_aa_types=20)
    g.ds(p + "/interaction_param", np.array([[[6.0, 1.0, -0.1, 1.0]]]))
    g.ds(p + "/aa_types", np.array([restype_order['PRO' if seq[i] == 'CPR' else seq[i]] for i in res_id]))
    g.ds(p + "/index1", np.arange(nd + na))
    g.ds(p + "/type1", np.arange(nd + na) * 0)
    g.ds(p + "/id1", res_id)
    g.ds(p + "/index2", np.arange(n_sc))
    g.ds(p + "/type2", 0 * np.arange(n_sc))
    g.ds(p + "/id2", sc_resnum)

    for name, i1, t1, i2, t2 in [("hb_environment_coverage_hn", np.arange(nd), donors, np.arange(nd, nd + na), acceptors),
                                 ("hb_environment_coverage_oc", np.arange(nd, nd + na), acceptors, np.arange(nd), donors)]:
        p = g.group(name, ["infer_H_O", "infer_H_O"])
        g.ds(p + "/interaction_param", np.array([[[3.5, 2.0, 0.0, 2.0]]]))
        g.ds(p + "/index1", i1)
        g.ds(p + "/type1", i1 * 0)
        g.ds(p + "/id1", t1)
        g.ds(p + "/index2", i2)
        g.ds(p + "/type2", i2 * 0)
        g.ds(p + "/id2", t2)

    p = g.group("cat_pos_bb_coverage", ["hb_environment_coverage_hn", "hb_environment_coverage_oc"], n_dim=1)
    g.ds(p + "/index_pos1", np.arange(nd))
    g.ds(p + "/index_pos2", np.arange(na))

    bb = params["bb_env"]
    p = g.group("bb_sigmoid_coupling_environment", ["environment_coverage_hb", "cat_pos_bb_coverage"],
                integrator_level=1, num_aa_types=20, scale=bb[0], center=bb[1], sharpness=bb[2],
                hbond_weight=bb[3])
    g.ds(p + "/weights", params["environment/weights"])


def _tail_nodes(g, n_res, sc, pl):
    """write_weighted_pos, write_rama_coord2, write_affine_alignment
    (py/upside_config.py:370-377,407-451,314-330)."""
    n_atom = 3 * n_res
    p = g.group("weighted_pos", [sc, pl])
    n_sc = len(g.dsets["/input/potential/%s/affine_residue" % sc])
    g.ds(p + "/index_pos", np.arange(n_sc))
    g.ds(p + "/index_weight", np.arange(n_sc))

    N = 3 * np.arange(n_res)
    id1 = np.column_stack((N - 1, N, N + 1, N + 2))
    id2 = np.column_stack((N, N + 1, N + 2, N + 3))
    id2[id2 >= n_atom] = -1
    id3 = np.column_stack((N - 1, N, N + 1, N + 2, N + 3))
    id3[id3 >= n_atom] = -1
    p = g.group("Dihedral_phi", ["pos"])
    g.ds(p + "/id", id1)
    p = g.group("Dihedral_psi", ["pos"])
    g.ds(p + "/id", id2)
    p = g.group("rama_coord", ["Dihedral_phi", "Dihedral_psi"])
    g.ds(p + "/index_pos1", np.arange(n_res))
    g.ds(p + "/index_pos2", np.arange(n_res))
    g.ds(p + "/id", id3)

    ref = np.zeros((n_res, 3, 3))
    ref[:] = _REF_NCAC_CB[:3]
    ref -= ref.mean(axis=1)[:, None]
    p = g.group("affine_alignment", ["pos"])
    g.ds(p + "/atoms", np.column_stack((N, N + 1, N + 2)))
    g.ds(p + "/ref_geom", ref)


def build_ff21(seq, pos, rama_pot=None, params=None, rama_seed=0, damping=0.4, max_iter=1000, tol=1e-3,
               normalize_rama=True):
    """Build the ff_2.1 `/input` graph for a single chain.

    seq: list of three-letter residue names; pos: (3*n_res, 3) backbone N,CA,C.
    rama_pot given with normalize_rama=False is stored as is (maps that are
    already normalised, e.g. taken from an existing `.up`).
    Node order follows upside_config.main (py/upside_config.py:2398-2536).
    """
    params = load_params() if params is None else params
    seq = list(seq)
    n_res = len(seq)
    g = Graph()
    g.ds("/input/sequence", np.array(seq, dtype="S3"))
    g.ds("/input/pos", np.asarray(pos, dtype=np.float32).reshape(3 * n_res, 3, 1))
    _spring_nodes(g, seq, 3 * n_res)
    if rama_pot is None:
        rama_pot = standin_rama_maps(seq, seed=rama_seed)
    _rama_nodes(g, seq, rama_pot, params, normalize=normalize_rama)
    _backbone_pairs(g, [('PRO' if s == 'CPR' else s) for s in seq])
    donors, acceptors = _hbond_nodes(g, [('PRO' if s == 'CPR' else s) for s in seq], params)
    sc, pl, beadtypes, sc_resnum, id_seq = _rotamer_placement(g, seq, params)
    _rotamer_backbone(g, seq, params, donors, acceptors, sc, beadtypes, sc_resnum)
    _rotamer(g, params, sc, pl, beadtypes, id_seq, damping=damping, max_iter=max_iter, tol=tol)
    _environment(g, seq, params, sc, sc_resnum)
    _bb_environment(g, seq, params, donors, acceptors, sc_resnum)
    _tail_nodes(g, n_res, sc, pl)
    _pivot_moves(g, n_res)
    return g


def _pivot_moves(g, n_res):
    """/input/pivot_moves for the Monte Carlo pivot sampler
    (py/upside_config.py:2543-2552): every residue with a full
    (prevC, N, CA, C, nextN) window, its Rama proposal layer, and the atoms
    after nextN as the rotated range."""
    n_atom = 3 * n_res
    ids = g.dsets["/input/potential/rama_coord/id"]
    keep = np.array([-1 not in tuple(x) for x in ids])
    atom = ids[keep].astype(np.int32)
    g.ds("/input/pivot_moves/proposal_pot", g.dsets["/input/potential/rama_map_pot/rama_pot"])
    g.ds("/input/pivot_moves/pivot_atom", atom)
    g.ds("/input/pivot_moves/pivot_restype", g.dsets["/input/potential/rama_map_pot/rama_map_id_all"][keep])
    g.ds("/input/pivot_moves/pivot_range",
         np.column_stack((atom[:, 4] + 1, np.zeros(len(atom), "i") + n_atom)).astype(np.int32))


def write_up(path, g):
    with h5lite.File(path, "w") as f:
        f.create_group("/input/potential")
        for p in g.order:
            f.create_group(p)
            for k, v in g.groups[p].items():
                f.set_attr(p, k, v)
        for p, arr in g.dsets.items():
            f.write(p, arr)


def make_chain(path, n_res, seed, rama_seed=0, **kw):
    """Seeded random-sequence chain with a random-Rama backbone, written to `path`."""
    rng = np.random.default_rng(seed)
    seq = random_sequence(n_res, rng)
    pos = random_initial_config(n_res, rng)
    g = build_ff21(seq, pos, rama_seed=rama_seed, **kw)
    write_up(path, g)
    return seq, pos
