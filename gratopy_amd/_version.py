__version__ = "0.2.0a2+mi355x"
