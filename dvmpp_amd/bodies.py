"""Closed, counter-clockwise body contours in the reference's body-file form
(rows x z, last point equal to the first; src/VortexSheet.cpp:56-102)."""
import numpy as np


def close(x, z):
    x = np.append(x, x[0])
    z = np.append(z, z[0])
    return np.ascontiguousarray(x), np.ascontiguousarray(z)


def cylinder(nl, radius=0.5):
    th = 2.0 * np.pi * np.arange(nl) / nl
    return close(radius * np.cos(th), radius * np.sin(th))


def ellipse(nl, aspect=2.0, aoa_deg=0.0, chord=1.0):
    """Ellipse of major axis `chord`, minor axis chord/aspect, rotated by -aoa
    (nose up for a free stream along +x)."""
    th = 2.0 * np.pi * np.arange(nl) / nl
    a, b = 0.5 * chord, 0.5 * chord / aspect
    x, z = a * np.cos(th), b * np.sin(th)
    al = -np.deg2rad(aoa_deg)
    return close(x * np.cos(al) - z * np.sin(al), x * np.sin(al) + z * np.cos(al))


def rounded_square(nl, side=1.0, radius=0.15):
    """Square of the given side with rounded corners, nl points spaced evenly
    in arc length."""
    h = 0.5 * side - radius
    # parametrise by arc length: 4 straight edges of length 2h, 4 quarter arcs
    edge, arc = 2.0 * h, 0.5 * np.pi * radius
    per = 4 * (edge + arc)
    s = per * np.arange(nl) / nl
    x = np.empty(nl)
    z = np.empty(nl)
    centres = [(h, h), (-h, h), (-h, -h), (h, -h)]
    for k in range(nl):
        t = s[k]
        for q in range(4):
            if t < edge:  # straight edge q, starting at angle q*90 deg
                ang = q * 0.5 * np.pi
                cx, cz = centres[q - 1] if q > 0 else centres[3]
                # edge from the end of arc q-1 to the start of arc q
                px = cx + radius * np.cos(ang)
                pz = cz + radius * np.sin(ang)
                tx, tz = -np.sin(ang), np.cos(ang)
                x[k], z[k] = px + t * tx, pz + t * tz
                break
            t -= edge
            if t < arc:
                cx, cz = centres[q]
                ang = q * 0.5 * np.pi + t / radius
                x[k], z[k] = cx + radius * np.cos(ang), cz + radius * np.sin(ang)
                break
            t -= arc
    return close(x, z)
