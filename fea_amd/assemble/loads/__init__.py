from .base import BaseLoad
from .pressure import Pressure
from .traction import Traction
from .contact import NodeToSurfaceContact
from .point_loads import Concentrate_Force, Moment
from .barrier import ContactSelf, Contact
