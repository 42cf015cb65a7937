"""The error raised for any problem in a recipe (recipe/errors.py of the reference); the CLI exits 2 on it."""


class RecipeError(Exception):
    """A recipe problem, with a message that names the key."""
