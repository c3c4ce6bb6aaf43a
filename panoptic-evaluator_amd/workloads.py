"""Seeded synthetic workloads of the benchmark configurations.

Voronoi panoptic maps and rotated-ellipse detection masks, generated with
torch on any device (GPU for full-size benchmarks, CPU for parity tests).
Every input comes from one seed; nothing is read from disk.

  config 0: 8 x 512 x 512, 133 classes (0..79 thing), 40 seeds
  config 1: 16 x 800 x 1088, 80 seeds, 100 ellipse detections per image
  config 2: 256 mixed sizes (multiples of 32 in [320, 1344]), 5..120 segments,
            0..300 detections
  config 3: 64 x 1024 x 2048, 19 classes (8 thing), 150 seeds
  config 4: serving requests of 1..64 images (256..1024 square), 10..100
            segments; per-image PQ or prefix rewards over 8..128 masks with
            checkpoints every prefix or every 8th
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch


def _rand(gen, *shape, device):
    return torch.rand(shape, generator=gen, device=device, dtype=torch.float64)


def _randint(gen, low, high, shape, device):
    return torch.randint(low, high, shape, generator=gen, device=device)


def nearest_seed(seeds: torch.Tensor, H: int, W: int, rows_per_block: int = 64) -> torch.Tensor:
    """Index of the nearest seed for every pixel (ties -> lower index)."""
    device = seeds.device
    out = torch.empty((H, W), dtype=torch.int64, device=device)
    xs = torch.arange(W, device=device, dtype=torch.float32)
    sy, sx = seeds[:, 0].float(), seeds[:, 1].float()
    for r0 in range(0, H, rows_per_block):
        ys = torch.arange(r0, min(H, r0 + rows_per_block), device=device, dtype=torch.float32)
        d = (ys[:, None, None] - sy) ** 2 + (xs[None, :, None] - sx) ** 2
        out[r0:r0 + len(ys)] = d.argmin(-1)
    return out


@dataclass
class PanopticPair:
    gt: torch.Tensor        # B x H x W int32 slots
    pred: torch.Tensor      # B x H x W int32 slots
    gc: torch.Tensor        # B x G int32
    pc: torch.Tensor        # B x P int32
    crowd: torch.Tensor     # B x G bool


def voronoi_image(gen, H, W, seeds, num_classes, num_things, *, void=0.10, crowd_p=0.05, jitter=8.0,
                  spurious=0.15, flip=0.10, stuff_density=1.0, device="cpu", capacity=None):
    """One image pair: target slots 1..seeds in seed order, prediction from
    jittered seeds plus spurious seeds.  Returns gt, pred (H x W int32) and
    metadata vectors (gc, pc, crowd) padded to ``capacity`` when given."""
    pos = _rand(gen, seeds, 2, device=device) * torch.tensor([H, W], device=device, dtype=torch.float64)
    classes = _randint(gen, 0, num_classes, (seeds,), device)
    if stuff_density != 1.0:
        # stuff seeds are thinned so that stuff regions are large
        keep = (classes < num_things) | (_rand(gen, seeds, device=device) < stuff_density)
        pos, classes = pos[keep], classes[keep]
        seeds = int(pos.shape[0])
    gt = nearest_seed(pos, H, W) + 1
    if void > 0:
        gt = torch.where(_rand(gen, H, W, device=device) < void, 0, gt)
    extra = int(round(spurious * seeds))
    jittered = pos + torch.randn((seeds, 2), generator=gen, device=device, dtype=torch.float64) * jitter
    jittered[:, 0].clamp_(0, H - 1)
    jittered[:, 1].clamp_(0, W - 1)
    extra_pos = _rand(gen, extra, 2, device=device) * torch.tensor([H, W], device=device, dtype=torch.float64)
    pred = nearest_seed(torch.cat([jittered, extra_pos]), H, W) + 1
    pclasses = torch.cat([classes, _randint(gen, 0, num_classes, (extra,), device)])
    flips = _rand(gen, seeds + extra, device=device) < flip
    pclasses = torch.where(flips, _randint(gen, 0, num_classes, (seeds + extra,), device), pclasses)
    crowd = (_rand(gen, seeds, device=device) < crowd_p) & (classes < num_things)
    G, P = seeds + 1, seeds + extra + 1
    gcap = capacity[0] if capacity else G
    pcap = capacity[1] if capacity else P
    ga = torch.bincount(gt.flatten(), minlength=G)
    pa = torch.bincount(pred.flatten(), minlength=P)
    gc = torch.full((gcap,), -1, dtype=torch.int32, device=device)
    pc = torch.full((pcap,), -1, dtype=torch.int32, device=device)
    cr = torch.zeros(gcap, dtype=torch.bool, device=device)
    gc[1:G] = torch.where(ga[1:] > 0, classes, -1).int()
    pc[1:P] = torch.where(pa[1:] > 0, pclasses, -1).int()
    cr[1:G] = crowd & (ga[1:] > 0)
    return gt.int(), pred.int(), gc, pc, cr


def panoptic_batch(seed, B, H, W, seeds, num_classes, num_things, device="cpu", **kwargs) -> PanopticPair:
    gen = torch.Generator(device=device).manual_seed(seed)
    extra = int(round(kwargs.get("spurious", 0.15) * seeds))
    capacity = (seeds + 1, seeds + extra + 1)
    parts = [voronoi_image(gen, H, W, seeds, num_classes, num_things, device=device, capacity=capacity,
                           **kwargs) for _ in range(B)]
    gt, pred, gc, pc, crowd = (torch.stack(x) for x in zip(*parts))
    return PanopticPair(gt.contiguous(), pred.contiguous(), gc, pc, crowd)


def ellipse_masks(gen, N, H, W, device="cpu", min_axis=8.0, max_axis=200.0, block=16) -> torch.Tensor:
    """N rotated ellipses (semi-axes U[min, max], centre uniform, angle
    U[0, pi)) as an N x H x W bool tensor; every mask has its centre pixel.
    Masks are evaluated ``block`` at a time (same float32 formula per pixel)."""
    centre = _rand(gen, N, 2, device=device) * torch.tensor([H, W], device=device, dtype=torch.float64)
    axes = min_axis + _rand(gen, N, 2, device=device) * (max_axis - min_axis)
    angle = _rand(gen, N, device=device) * math.pi
    ys = torch.arange(H, device=device, dtype=torch.float32)[None, :, None]
    xs = torch.arange(W, device=device, dtype=torch.float32)[None, None, :]
    masks = torch.empty((N, H, W), dtype=torch.bool, device=device)
    cy, cx = centre[:, 0].float(), centre[:, 1].float()
    a, b = axes[:, 0].float(), axes[:, 1].float()
    cos, sin = torch.cos(angle).float(), torch.sin(angle).float()
    for n0 in range(0, N, block):
        sl = slice(n0, min(N, n0 + block))
        col = lambda t: t[sl][:, None, None]  # noqa: E731
        dy, dx = ys - col(cy), xs - col(cx)
        u = (dx * col(cos) + dy * col(sin)) / col(a)
        v = (-dx * col(sin) + dy * col(cos)) / col(b)
        masks[sl] = u * u + v * v <= 1.0
    rows = cy.clamp(0, H - 1).long()
    cols = cx.clamp(0, W - 1).long()
    masks[torch.arange(N, device=device), rows, cols] = True
    return masks


def config0(device="cpu", seed=0, B=8, H=512, W=512):
    return panoptic_batch(seed, B, H, W, 40, 133, 80, device=device)


def config1(device="cpu", seed=1, B=16, H=800, W=1088, detections=100):
    pair = panoptic_batch(seed, B, H, W, 80, 133, 80, device=device)
    gen = torch.Generator(device=device).manual_seed(seed + 1000)
    masks = [ellipse_masks(gen, detections, H, W, device=device) for _ in range(B)]
    classes = [_randint(gen, 0, 80, (detections,), device) for _ in range(B)]
    scores = [_rand(gen, detections, device=device) for _ in range(B)]
    return pair, masks, classes, scores


@dataclass
class UnevenBatch:
    """Per-image lists of a batch of uneven images (config 2)."""

    sizes: list          # (H, W)
    gt: list             # H x W int32
    pred: list
    gc: list             # S_i int32
    pc: list
    crowd: list          # S_i bool
    masks: list          # D_i x H x W bool
    classes: list        # D_i int64
    scores: list         # D_i float64


def config2(device="cpu", seed=2, B=256, sizes=(10, 42), segments=(5, 120), detections=(0, 300),
            num_classes=133, num_things=80):
    """B images with H and W drawn from multiples of 32 in [32 * sizes[0],
    32 * sizes[1]], U[segments] Voronoi seeds and U[detections] ellipses."""
    gen = torch.Generator(device=device).manual_seed(seed)
    out = UnevenBatch([], [], [], [], [], [], [], [], [])
    for _ in range(B):
        hw = _randint(gen, sizes[0], sizes[1] + 1, (2,), device)
        H, W = 32 * int(hw[0]), 32 * int(hw[1])
        seeds = int(_randint(gen, segments[0], segments[1] + 1, (1,), device))
        D = int(_randint(gen, detections[0], detections[1] + 1, (1,), device))
        gt, pred, gc, pc, crowd = voronoi_image(gen, H, W, seeds, num_classes, num_things, device=device)
        out.sizes.append((H, W))
        out.gt.append(gt)
        out.pred.append(pred)
        out.gc.append(gc)
        out.pc.append(pc)
        out.crowd.append(crowd)
        out.masks.append(ellipse_masks(gen, D, H, W, device=device))
        out.classes.append(_randint(gen, 0, num_things, (D,), device))
        out.scores.append(_rand(gen, D, device=device))
    return out


@dataclass
class RewardImage:
    size: tuple
    gt: torch.Tensor
    pred: torch.Tensor
    gc: torch.Tensor
    pc: torch.Tensor
    crowd: torch.Tensor
    masks: torch.Tensor | None = None     # prefix requests: K x H x W generated masks
    classes: torch.Tensor | None = None


def config4(device="cpu", seed=4, pool=16, images=(1, 64), sides=(256, 512, 768, 1024), segments=(10, 100),
            masks=(8, 128), strides=(1, 8), num_classes=133, num_things=80):
    """Serving requests: ``pool`` per-image-PQ requests and ``pool`` prefix
    requests, each with U[images] images of a side drawn from ``sides`` and
    U[segments] Voronoi seeds; prefix requests generate U[masks] ellipse masks
    per image and evaluate every prefix or every 8th (the last always)."""
    gen = torch.Generator(device=device).manual_seed(seed)
    pq, prefix = [], []
    for kind in ("pq", "prefix"):
        for _ in range(pool):
            count = int(_randint(gen, images[0], images[1] + 1, (1,), device))
            K = int(_randint(gen, masks[0], masks[1] + 1, (1,), device))
            stride = strides[int(_randint(gen, 0, len(strides), (1,), device))]
            request = []
            for _ in range(count):
                side = sides[int(_randint(gen, 0, len(sides), (1,), device))]
                seeds = int(_randint(gen, segments[0], segments[1] + 1, (1,), device))
                gt, pred, gc, pc, crowd = voronoi_image(gen, side, side, seeds, num_classes, num_things,
                                                        device=device)
                image = RewardImage((side, side), gt, pred, gc, pc, crowd)
                if kind == "prefix":
                    image.masks = ellipse_masks(gen, K, side, side, device=device)
                    image.classes = _randint(gen, 0, num_classes, (K,), device).int()
                request.append(image)
            if kind == "pq":
                pq.append(request)
            else:
                points = None if stride == 1 else sorted(set(list(range(stride, K + 1, stride)) + [K]))
                prefix.append((request, points))
    return pq, prefix


def config3(device="cpu", seed=3, B=64, H=1024, W=2048):
    return panoptic_batch(seed, B, H, W, 150, 19, 8, device=device, void=0.02, jitter=4.0, flip=0.05,
                          stuff_density=0.2)


def isthing_vector(num_classes, num_things):
    return [c < num_things for c in range(num_classes)]


def rle_encode(masks: torch.Tensor) -> list:
    """COCO RLE dicts of N x H x W bool masks in pycocotools' compressed-string
    form (synthetic inputs for the RLE configuration): column-major runs are
    found on the masks' device, the strings are written on the host."""
    N, H, W = masks.shape
    flat = masks.transpose(1, 2).reshape(N, H * W)  # pixel (y, x) at x * H + y
    previous = torch.cat([torch.zeros((N, 1), dtype=torch.bool, device=masks.device), flat[:, :-1]], dim=1)
    rows, positions = (flat != previous).nonzero(as_tuple=True)
    rows, positions = rows.cpu().tolist(), positions.cpu().tolist()
    edges = [[0] for _ in range(N)]
    for r, p in zip(rows, positions):
        edges[r].append(p)
    out = []
    for n in range(N):
        e = edges[n] + [H * W]
        counts = [e[i + 1] - e[i] for i in range(len(e) - 1)]
        chars = []
        for i, count in enumerate(counts):
            x = count - counts[i - 2] if i > 2 else count
            while True:
                c = x & 0x1F
                x >>= 5
                more = (x != -1) if (c & 0x10) else (x != 0)
                chars.append(chr(48 + (c | 0x20 if more else c)))
                if not more:
                    break
        out.append({"size": [H, W], "counts": "".join(chars).encode("ascii")})
    return out
