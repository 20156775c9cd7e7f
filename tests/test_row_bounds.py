"""The brick kernel's row bound (bound_rows in rexlib_amd/csrc/insertion.hip),
restated in float32 numpy: the per-axis bound of
src/backends/cpu/kernels/row_interval.cpp:34-136 over the whole k_x range,
then one refinement through the integer k_x that can reach the box from
those rows. Against brute force over random planes and 8^3 bricks, the
refined bound keeps every row that holds a sample whose first tap lies in
the brick, and drops about half the rows of the unrefined one."""
import numpy as np

f32 = np.float32


def rotation(rng):
    q = rng.normal(size=4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def reciprocal(x):
    return f32(0) if x == 0 else f32(1) / f32(x)


def bound_rows(xs, ys, lower, upper, sign_index, refine, n=256, last_shell=128):
    h = n // 2 + 1
    widest = min(h - 2, last_shell)
    first_row, last_row = max(-(n // 2 - 1), -widest), min(n // 2 - 1, widest)
    sign = f32(1 if sign_index == 0 else -1)
    first = sign_index
    low, high = f32(first_row), f32(last_row)
    inv_y = [reciprocal(y) for y in ys]
    inv_x = [reciprocal(x) for x in xs]
    for a in range(3):
        step, y_step = sign * xs[a], sign * ys[a]
        at_first, at_last = f32(first) * step, f32(widest) * step
        lo, hi = f32(lower[a] - max(at_first, at_last)), f32(upper[a] - min(at_first, at_last))
        if y_step == 0:
            if not lo <= 0 <= hi:
                return 0, 0
        else:
            l, u = f32(lo * sign * inv_y[a]), f32(hi * sign * inv_y[a])
            low, high = max(low, min(l, u)), min(high, max(l, u))
    if low > high:
        return 0, 0
    if refine:
        k_low, k_high = f32(first), f32(widest)
        for a in range(3):
            step = sign * xs[a]
            if step != 0:
                y_step = sign * ys[a]
                b0, b1 = f32(low * y_step), f32(high * y_step)
                inv = sign * inv_x[a]
                l = f32(f32(lower[a] - max(b0, b1)) * inv)
                u = f32(f32(upper[a] - min(b0, b1)) * inv)
                k_low, k_high = max(k_low, min(l, u)), min(k_high, max(l, u))
        k_low, k_high = np.ceil(k_low), np.floor(k_high)
        if k_low > k_high:
            return 0, 0
        for a in range(3):
            step, y_step = sign * xs[a], sign * ys[a]
            if y_step != 0:
                at_first, at_last = f32(k_low * step), f32(k_high * step)
                inv = sign * inv_y[a]
                l = f32(f32(lower[a] - max(at_first, at_last)) * inv)
                u = f32(f32(upper[a] - min(at_first, at_last)) * inv)
                low, high = max(low, min(l, u)), min(high, max(l, u))
        if low > high:
            return 0, 0
    return max(int(np.ceil(low)), first_row), min(int(np.floor(high)) + 1, last_row + 1)


def test_refined_bound_keeps_every_row_with_a_sample():
    rng = np.random.default_rng(5)
    rows_plain = rows_refined = rows_with_samples = 0
    for _ in range(600):
        r = rotation(rng)
        xs, ys = r[0].astype(f32), r[1].astype(f32)
        centre = rng.uniform(-127, 127, 3)
        normal = np.cross(r[0], r[1])
        centre = centre - np.dot(centre, normal) * normal + rng.uniform(-2, 2) * normal
        lo = np.floor(centre) - 4
        lower, upper = (lo - 0.01).astype(f32), (lo + 8 + 0.01).astype(f32)
        for sign_index in (0, 1):
            sign = 1 if sign_index == 0 else -1
            p0, p1 = bound_rows(xs, ys, lower, upper, sign_index, refine=False)
            r0, r1 = bound_rows(xs, ys, lower, upper, sign_index, refine=True)
            rows_plain += max(p1 - p0, 0)
            rows_refined += max(r1 - r0, 0)
            ky = np.arange(-127, 128)[:, None]
            kx = np.arange(sign_index, 128)[None, :]
            # Positions as the kernel forms them: k_y Y in double rounded to
            # float, plus k_x X in float.
            base = (ky.astype(np.float64)[..., None] * (sign * ys.astype(np.float64))).astype(f32)
            position = base + (kx.astype(f32)[..., None] * (sign * xs)).astype(f32)
            cell = np.floor(position)
            inside = np.all((cell >= lo) & (cell < lo + 8), axis=-1) & (kx * kx + ky * ky <= 128.5 ** 2)
            rows = np.unique(np.nonzero(inside)[0] - 127)
            rows_with_samples += rows.size
            assert np.all((rows >= r0) & (rows < r1)), (rows, r0, r1)
    assert rows_with_samples > 1000
    assert rows_refined < 0.6 * rows_plain
