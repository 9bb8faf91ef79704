"""CycleError as the reference defines it (sparse/csgraph/cycleerror.py:1-6):
a ValueError whose constructor forwards (args, kwargs) as one tuple, so
str(e) is that tuple's repr, and whose ``vertex`` attribute is required."""


class CycleError(ValueError):

    def __init__(self, *args, **kwargs):
        super(CycleError, self).__init__(args, kwargs)
        self.vertex = kwargs['vertex']
