"""ORACLE — test infrastructure only; the product never imports this module.

Restatement of how Xyce turns per-device Jacobian stamps into the fixed CSR
pattern and per-instance offsets:

* ``jac_stamp_map``  — DeviceInstance::jacStampMap
  (src/DeviceModelPKG/Core/src/N_DEV_DeviceInstance.C:663-883), which merges
  an eliminated internal node (d' or s' when RD/RS = 0) into the node it
  collapses onto, producing the reduced stamp plus the ``map``/``map2``
  tables that registerJacLIDs uses.
* ``mos1_stamps`` — the four MOSFET1 old-Meyer stamps built in the instance
  constructor (N_DEV_MOSFET1.C:1123-1168) and ``jacobianStamp``
  (N_DEV_MOSFET1.C:1579-1589).
* ``build_pattern`` — TopoLSUtil::generateRowColData stamp branch
  (src/TopoManagerPKG/src/N_TOP_TopoLSUtil.C:2375-2430: per-row column
  lists, ``sort`` + ``unique`` over GIDs, ground GID -1), followed by the
  Epetra fill that re-indexes columns by local id with ground last
  (src/LinearAlgebraServicesPKG/src/N_LAS_Builder.C:336-338, 403-445).
* ``stamp_offsets`` — Indexor::matrixGlobalToLocal
  (src/TopoManagerPKG/src/N_TOP_Indexor.C:171-219): each stamp entry becomes
  its position inside the sorted CSR row.
* ``mos1_register_offsets`` — MOSFET1 Instance::registerJacLIDs
  (N_DEV_MOSFET1.C:1606-1673): the 22 named offsets, read through map/map2.
"""

import copy


def jac_stamp_map(stamp_parent, map_parent, map2_parent, frm, to, original_size):
    """Literal restatement of DeviceInstance::jacStampMap.

    Returns (stamp, map, map2).  ``map_parent``/``map2_parent`` may be empty
    lists, in which case identity maps are made (N_DEV_DeviceInstance.C:687-699).
    """
    if frm <= to:
        raise ValueError("jacStampMap: from <= to")
    stamp_parent = copy.deepcopy(stamp_parent)
    map_parent = list(map_parent)
    map2_parent = copy.deepcopy(map2_parent)
    if len(map_parent) == 0:
        map_parent = list(range(original_size))
        map2_parent = [list(range(len(stamp_parent[i]))) for i in range(original_size)]
    map2 = [None] * original_size

    for i in range(original_size):
        f_index = -1
        t_index = -1
        p_row = map_parent[i]
        for j in range(len(stamp_parent[p_row])):
            if stamp_parent[p_row][j] == frm:
                f_index = j
            if stamp_parent[p_row][j] == to:
                t_index = j
        map2[i] = [0] * len(map2_parent[i])
        f_index2 = -1
        t_index2 = -1
        for j in range(len(map2_parent[i])):
            map2[i][j] = map2_parent[i][j]
            if stamp_parent[p_row][map2[i][j]] == frm:
                f_index2 = j
            if stamp_parent[p_row][map2[i][j]] == to:
                t_index2 = j
        if f_index >= 0:
            if t_index >= 0:
                for j in range(len(map2[i])):
                    if map2[i][j] > f_index:
                        map2[i][j] -= 1
                if f_index2 >= 0 and t_index2 >= 0:
                    map2[i][f_index2] = map2[i][t_index2]
            else:
                for j in range(len(map2[i])):
                    v = stamp_parent[p_row][map2[i][j]]
                    if to < v < frm:
                        map2[i][j] += 1
                t_index = 0
                for j in range(len(stamp_parent[p_row])):
                    if to > stamp_parent[p_row][j]:
                        t_index += 1
                    else:
                        break
                if f_index2 >= 0:
                    map2[i][f_index2] = t_index

    mp = [0] * original_size
    p_size = len(stamp_parent)

    map2_tmp = copy.deepcopy(map2)
    i = 0
    while i < len(stamp_parent[frm]) and stamp_parent[frm][i] < p_size - 1:
        new_col = False
        for j in range(len(stamp_parent[to])):
            if j == 0:
                new_col = True
            if stamp_parent[frm][i] == stamp_parent[to][j]:
                new_col = False
                break
        if new_col:
            for j in range(len(map2[to])):
                if stamp_parent[to][map2_tmp[to][j]] > stamp_parent[frm][i]:
                    map2[to][j] += 1
        i += 1

    dup = [0] * original_size
    f_mod = frm
    for i in range(1, original_size):
        dup[i] = -1
        for j in range(i):
            if map_parent[i] == map_parent[j]:
                dup[i] = j
                if i <= f_mod:
                    f_mod += 1
                break

    for i in range(f_mod):
        mp[i] = map_parent[i]
    mp[f_mod] = mp[to]
    for i in range(f_mod + 1, original_size):
        mp[i] = map_parent[i] - 1
    for i in range(1, original_size):
        if dup[i] >= 0:
            mp[i] = mp[dup[i]]

    map2_tmp = copy.deepcopy(map2)
    map_2_from = frm
    if mp[frm] != mp[to]:
        for i in range(frm + 1, original_size):
            if mp[i] == mp[to]:
                map_2_from = i
                break
        if map_2_from == frm:
            raise RuntimeError("jacStampMap internal error 2")

    i = 0
    while i < len(stamp_parent[to]) and stamp_parent[to][i] < p_size - 1:
        new_col = False
        for j in range(len(stamp_parent[frm])):
            if j == 0:
                new_col = True
            if stamp_parent[to][i] == stamp_parent[frm][j]:
                new_col = False
                break
        if new_col:
            for j in range(len(map2[map_2_from])):
                if stamp_parent[frm][map2_tmp[map_2_from][j]] > stamp_parent[to][i]:
                    map2[map_2_from][j] += 1
        i += 1

    fill = [[0] * p_size for _ in range(p_size)]
    for i in range(p_size):
        for c in stamp_parent[i]:
            fill[i][c] = 1
    for i in range(p_size):
        fill[to][i] += fill[frm][i]
    for i in range(p_size):
        fill[i][to] += fill[i][frm]
    for i in range(frm, p_size - 1):
        for j in range(p_size):
            fill[i][j] = fill[i + 1][j]
        for j in range(p_size):
            fill[j][i] = fill[j][i + 1]
    stamp = []
    for i in range(p_size - 1):
        stamp.append([j for j in range(p_size - 1) if fill[i][j] > 0])
    return stamp, mp, map2


# MOSFET1 old-Meyer full stamp, N_DEV_MOSFET1.C:1127-1157.
# Local node order: 0=d 1=g 2=s 3=b 4=d' 5=s'.
_MOS1_DC_SC = [
    [0, 4],
    [1, 3, 4, 5],
    [2, 5],
    [1, 3, 4, 5],
    [0, 1, 3, 4, 5],
    [1, 2, 3, 4, 5],
]


def mos1_stamps():
    """The four (stamp, map, map2) triples of N_DEV_MOSFET1.C:1158-1166."""
    dc_sc = copy.deepcopy(_MOS1_DC_SC)
    map_dc_sc, map2_dc_sc = [], []
    # jacStampMap fills map_parent/map2_parent in place when empty; restate that.
    map_dc_sc = list(range(6))
    map2_dc_sc = [list(range(len(dc_sc[i]))) for i in range(6)]
    st_dc, map_dc, map2_dc = jac_stamp_map(dc_sc, map_dc_sc, map2_dc_sc, 5, 2, 6)
    st_sc, map_sc, map2_sc = jac_stamp_map(dc_sc, map_dc_sc, map2_dc_sc, 4, 0, 6)
    st, mp, map2 = jac_stamp_map(st_dc, map_dc, map2_dc, 4, 0, 6)
    return {
        (True, True): (dc_sc, map_dc_sc, map2_dc_sc),
        (True, False): (st_dc, map_dc, map2_dc),
        (False, True): (st_sc, map_sc, map2_sc),
        (False, False): (st, mp, map2),
    }


def mos1_gids(lids, has_dc, has_sc):
    """Instance GID vector in stamp order: external d,g,s,b then internal d', s'
    when present (CktGraphBasic::registerJacLIDswithDevs ext-then-int order,
    N_TOP_CktGraphBasic.C:1116-1119; registerLIDs N_DEV_MOSFET1.C:1320-1332)."""
    g = list(lids[:4])
    if has_dc:
        g.append(lids[4])
    if has_sc:
        g.append(lids[5])
    return g


def build_pattern(n_rows, device_rows):
    """CSR over rows 0..n_rows (row n_rows = ground) from device stamps.

    ``device_rows`` yields (gids, stamp) per device in node-list order; gid -1
    is ground.  Returns (rowptr, colidx) with local ids, ground = n_rows.
    """
    cols = [[] for _ in range(n_rows + 1)]
    gl = lambda g: n_rows if g == -1 else g
    for gids, stamp in device_rows:
        for i, row in enumerate(stamp):
            for c in row:
                if c < len(gids):
                    cols[gl(gids[i])].append(gids[c])
    rowptr = [0]
    colidx = []
    for r in range(n_rows + 1):
        # std::list::sort + unique over GIDs (ground -1 sorts first) ...
        uniq = sorted(set(cols[r]))
        # ... then Epetra's local column ids, which put ground (GID -1,
        # last slot of SOLUTION_OVERLAP_GND) after every other column.
        loc = sorted(gl(g) for g in uniq)
        colidx.extend(loc)
        rowptr.append(len(colidx))
    return rowptr, colidx


def stamp_offsets(rowptr, colidx, n_rows, gids, stamp, rowmaps=None):
    """Indexor::matrixGlobalToLocal: stamp entry -> offset in its CSR row.
    ``rowmaps`` (a dict owned by the caller, one per pattern) caches each
    row's column->offset map, as the reference's accelerated matrix index
    does."""
    gl = lambda g: n_rows if g == -1 else g
    if rowmaps is None:
        rowmaps = {}
    out = []
    for i, row in enumerate(stamp):
        r = gl(gids[i])
        pos = rowmaps.get(r)
        if pos is None:
            rc = colidx[rowptr[r]:rowptr[r + 1]]
            pos = rowmaps[r] = {c: k for k, c in enumerate(rc)}
        out.append([pos[gl(gids[c])] for c in row])
    return out


# Order of the 22 offsets assigned by registerJacLIDs, as (row, map2 column)
# pairs into the *unmerged* 6-row stamp (N_DEV_MOSFET1.C:1650-1673).
MOS1_OFFSET_ORDER = [
    (0, 0), (0, 1),
    (1, 0), (1, 1), (1, 2), (1, 3),
    (2, 0), (2, 1),
    (3, 0), (3, 1), (3, 2), (3, 3),
    (4, 0), (4, 1), (4, 2), (4, 3), (4, 4),
    (5, 0), (5, 1), (5, 2), (5, 3), (5, 4),
]


def mos1_register_offsets(jac_lid_vec, mp, map2):
    """Instance::registerJacLIDs: offset_k = jacLIDVec[map[r]][map2[r][c]]."""
    return [jac_lid_vec[mp[r]][map2[r][c]] for (r, c) in MOS1_OFFSET_ORDER]
