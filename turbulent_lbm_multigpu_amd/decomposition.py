"""Domain decomposition geometry: CDomain, CComm and CManager::initSimulation.

Restates the reference's 3-D block decomposition (src/CManager.hpp:49-204,
src/CComm.hpp:8-79, src/CDomain.hpp:17-57): subdomain `id` sits at
(nx, ny, nz) = (id % Nx, id / Nx % Ny, id / (Nx*Ny)); faces on the outer
boundary are OBSTACLE, internal faces are GHOST with a CComm to the
neighbour; neighbouring subdomains overlap by two cells.  A periodic axis
(new) turns the outer faces of that axis into GHOST faces whose neighbour
wraps around.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

from ._capi import FLAG_GHOST_LAYER, FLAG_OBSTACLE

Vec3 = Tuple[int, int, int]


@dataclass
class CDomain:
    """src/CDomain.hpp:17-57"""
    uid: int
    size: Vec3
    origin: Vec3 = (0, 0, 0)
    length: Tuple[float, float, float] = (0.05, 0.05, 0.05)

    def getSize(self):
        return self.size

    def getOrigin(self):
        return self.origin

    def getLength(self):
        return self.length

    def getUid(self):
        return self.uid


@dataclass
class CComm:
    """src/CComm.hpp:8-79: one face exchange seen from the receiving subdomain."""
    dst_id: int
    send_size: Vec3
    recv_size: Vec3
    send_origin: Vec3
    recv_origin: Vec3
    comm_direction: Vec3


@dataclass
class Subdomain:
    uid: int
    coords: Vec3
    domain: CDomain
    bc: List[List[int]]            # BC[3][2]
    comms: List[CComm] = field(default_factory=list)
    has_lid: bool = False          # top row gets setGeometry() (src/CManager.hpp:200-202)


def subdomain_size(domain_size: Sequence[int], nums: Sequence[int]) -> Vec3:
    """CManager::setSubdomainNums (src/CManager.hpp:49-76)."""
    for a in range(3):
        if nums[a] <= 0 or domain_size[a] % nums[a] != 0:
            raise ValueError("Number of subdomains does not match with the grid size!")
    return tuple(int(domain_size[a] // nums[a]) for a in range(3))


def decompose(domain_size: Sequence[int], nums: Sequence[int], length=(0.1, 0.1, 0.1),
              periodic: Sequence[bool] = (False, False, False)) -> List[Subdomain]:
    """CManager::initSimulation for every rank (src/CManager.hpp:78-204)."""
    s = subdomain_size(domain_size, nums)
    sub_len = tuple(float(length[a]) / nums[a] for a in range(3))
    n_total = nums[0] * nums[1] * nums[2]
    out = []
    for uid in range(n_total):
        t = uid
        nx = t % nums[0]
        t //= nums[0]
        ny = t % nums[1]
        t //= nums[1]
        nz = t
        c = (nx, ny, nz)
        origin = (nx * s[0], ny * s[1], nz * s[2])
        bc = [[FLAG_GHOST_LAYER, FLAG_GHOST_LAYER] for _ in range(3)]
        for a in range(3):
            if not periodic[a]:
                if c[a] == 0:
                    bc[a][0] = FLAG_OBSTACLE
                if c[a] == nums[a] - 1:
                    bc[a][1] = FLAG_OBSTACLE
        sub = Subdomain(uid, c, CDomain(uid, s, origin, sub_len), bc)
        stride = (1, nums[0], nums[0] * nums[1])
        for a in range(3):
            for side in (0, 1):
                if bc[a][side] != FLAG_GHOST_LAYER:
                    continue
                nc = list(c)
                nc[a] = (c[a] + (1 if side else -1)) % nums[a]
                dst = nc[0] + nums[0] * (nc[1] + nums[1] * nc[2])
                size = [s[0], s[1], s[2]]
                size[a] = 1
                send_o = [0, 0, 0]
                recv_o = [0, 0, 0]
                normal = [0, 0, 0]
                if side == 0:
                    send_o[a], recv_o[a], normal[a] = 1, 0, 1
                else:
                    send_o[a], recv_o[a], normal[a] = s[a] - 2, s[a] - 1, -1
                sub.comms.append(CComm(dst, tuple(size), tuple(size), tuple(send_o), tuple(recv_o), tuple(normal)))
        sub.has_lid = ny == nums[1] - 1
        out.append(sub)
    return out


def lid_region(sub_size: Sequence[int]):
    """CController::setGeometry (src/CController.hpp:531-546): the
    VELOCITY_INJECTION layer at local y = sy-2, x in [1, sx-2], z in [1, sz-2]."""
    origin = (1, sub_size[1] - 2, 1)
    size = (sub_size[0] - 2, 1, sub_size[2] - 2)
    return origin, size


def monolithic_size(domain_size: Sequence[int], nums: Sequence[int]) -> Vec3:
    """Size of the single domain equivalent to a decomposition
    (src/main.cpp:358-361): N - 2(k - 1) per axis."""
    return tuple(int(domain_size[a] - 2 * (nums[a] - 1)) for a in range(3))


def key_origin(sub: Subdomain) -> Vec3:
    """Origin of a subdomain in the monolithic grid: n * (s - 2) per axis
    (neighbours overlap by two cells)."""
    s = sub.domain.size
    return tuple(sub.coords[a] * (s[a] - 2) for a in range(3))
