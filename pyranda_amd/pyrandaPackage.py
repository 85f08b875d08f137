"""Package base class (pyranda/pyrandaPackage.py:13-30)."""


class pyrandaPackage:

    def __init__(self, name, pyranda):
        self.name = name
        self.pyranda = pyranda

    def get_sMap(self):
        self.sMap = {}
