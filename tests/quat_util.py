"""Quaternion helpers of the tests (binary64): Euler degrees to the quaternion of
the same rotation as the Model rebuild's Rx Ry Rz (code/nans.cpp:1913-1924), the
Hamilton product, and rotation matrices."""
import numpy as np


def euler_to_quat(angles_deg):
    """[n][3] degrees -> [n][4] (w, x, y, z), q = qx(ax) qy(ay) qz(az)."""
    h = np.radians(np.asarray(angles_deg, np.float64)) / 2.0
    cx, cy, cz = np.cos(h[:, 0]), np.cos(h[:, 1]), np.cos(h[:, 2])
    sx, sy, sz = np.sin(h[:, 0]), np.sin(h[:, 1]), np.sin(h[:, 2])
    return np.stack([cx * cy * cz - sx * sy * sz, sx * cy * cz + cx * sy * sz,
                     cx * sy * cz - sx * cy * sz, cx * cy * sz + sx * sy * cz], axis=1)


def qmul(a, b):
    w1, v1 = a[0], np.asarray(a[1:])
    w2, v2 = b[0], np.asarray(b[1:])
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def axis_angle(axis, angle):
    """Rotation matrix of angle (rad) about a unit axis (Rodrigues)."""
    k = np.asarray(axis, np.float64)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def euler_matrix(angles_deg):
    """Rx Ry Rz of one body's Euler degrees."""
    ax, ay, az = np.radians(np.asarray(angles_deg, np.float64))
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    Rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


CORNERS = np.array([(0.5, 0.5, 0.5), (0.5, 0.5, -0.5), (-0.5, 0.5, 0.5), (-0.5, 0.5, -0.5),
                    (0.5, -0.5, 0.5), (0.5, -0.5, -0.5), (-0.5, -0.5, 0.5), (-0.5, -0.5, -0.5)])
