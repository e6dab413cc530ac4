"""Plane = point + normal (pyBrown/plane.py:30-57); defines the flux plane."""
import numpy as np


class Plane:
    def __init__(self, plane_point, normal_vector, label="XXX"):
        self.plane_point = np.array(plane_point, dtype=np.float64)
        self.normal_vector = np.array(normal_vector, dtype=np.float64)
        self.label = label

    def __str__(self):
        return "{}: plane_point = {}, normal_vector = {}".format(self.label, self.plane_point, self.normal_vector)

    __repr__ = __str__
